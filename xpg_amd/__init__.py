"""MI355X-native shoebox integration for XPG: the call surface of xpg/integrate.py, xpg/integrate_kernels.py and
xpg/shoebox.py over hand-written HIP kernels for gfx950 (libxpg_integrate.so, C ABI in include/xpg_integrate.h)."""
__version__ = "0.1.0"
