"""shearcuda_amd: MI355X-native (gfx950) nonsubsampled shearlet transform.

The product is the HIP library ``libshx.so`` (C ABI in include/shx.h); this
package holds its build script, the ctypes binding and the host-side mirror of
ShearCuda's MATLAB entry points (``api``).  Importing ``api`` requires the
built library and fails loudly without it.
"""
__version__ = "0.1.0"
