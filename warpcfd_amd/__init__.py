"""warpcfd_amd: MI355X-native (gfx950 HIP) SIMPLE/PISO inner loop with the
warp_cfd API (JohnCSu/WarpCFD).  See include/wcfd.h for the C ABI."""
__version__ = '0.1.0'
