"""CPU restatement of lightsmith's FDTD time stepping: TEST INFRASTRUCTURE ONLY.

This package is the checker that the HIP path is compared against. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only
as the checker or the timed CPU baseline -- never as the thing measured or shipped. The product
package ``lightsmith_amd`` does not import it and fails loudly when its HIP library is missing.

Modules
-------
fdtd     FP64 (or FP32) NumPy restatement of ``Engine`` (Yee curls, PEC, CPML, periodic/Bloch
         ghosts, pure and mixed ADE groups, off-diagonal tensor smoothing, DFT/time monitors).
sources  Point dipole, 1D-fed TFSF plane wave, table-fed TFSF (per-step read and scatter),
         Bloch plane source, mode source (per-step injection).
media    Poles, media, the discrete ADE permittivity, the dispersive time-step limit.
mie      Bohren & Huffman's Mie series of a homogeneous sphere (the scattering driver's whole-solve check).
tmm      Characteristic-matrix reflectance and transmittance of coherent layer stacks (the layer-stack and
         periodic drivers' whole-solve check).
rcwa     Rigorous coupled-wave analysis of lamellar gratings (the grating orders' whole-solve check).
dipole_substrate
         A small sphere above a substrate as a self-consistent point dipole with Sommerfeld reflection (the
         scattering driver's whole-solve check over a substrate).
cyee     A plain-C restatement of the bulk Yee + CPML + pure-ADE step (ctypes), for large
         sizes and the multi-core CPU baseline.

Every function cites the reference file:line it follows; paths are relative to
``src/lightsmith/backends/_fdtd/`` unless given in full. What pins each function is listed in
``tests/test_oracle_pins.py`` (closed forms, invariants and brute-force checks).
"""
