"""rumd_amd: MI355X-native implementation of RUMD's per-timestep hot path.

The compute path is hand-written gfx950 HIP behind the C ABI in include/rumdx.h
(librumdx.so).  The C++ host classes mirroring RUMD's Sample / Potential /
Integrator API and their Python surface live in the _rumd extension; importing
this package loads it and fails loudly when it has not been built.
"""
