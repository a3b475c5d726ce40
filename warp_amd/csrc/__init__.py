"""HIP sources of the gfx950 library and its build script."""
