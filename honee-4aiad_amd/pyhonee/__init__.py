"""Python helpers around the MI355X libCEED backend: mesh/tabulation generation, the ctypes binding of the
PETSc-free Newtonian driver (libhonee_setup.so) and reference-compatible file I/O.  The compute path is the native
library; these helpers only prepare inputs and move arrays."""
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
