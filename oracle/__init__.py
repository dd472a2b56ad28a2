"""CPU oracle for the GEOSmie Mie/LUT hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``geosmie_amd``)
imports, links or executes anything under ``oracle/``; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker (or as the timed CPU baseline).

Modules
-------
mie        numpy/scipy FP64 restatement of pymiecoated (homogeneous and
           coated coefficients, efficiencies, angle functions, S1/S2).
lut        numpy restatement of the runoptics size grid, lognormal PSD,
           Mueller elements, ``integratePSD`` and the post-integration block.
gsf        numpy restatement of ``spher_expan.f`` (Gauss nodes, linear
           interpolation, generalised spherical functions, expansion) and
           of the ``hbleg`` Legendre evaluator.
mpmie      arbitrary-precision Mie (Bohren & Huffman) on mpmath, an
           independent check of the FP64 restatement at any size.
runoptics  the Mie table loop of runoptics (every (bin, wavelength, rh)
           entry of one particle table) over the pieces above.
cmie       ctypes loader for ``liboracle_mie.so`` built from ``cmie.c``: the
           same homogeneous/coated algorithm in C (OpenMP), used as the fast
           oracle for large parity sizes and as the host-CPU baseline.

Each function cites the reference file:line it follows.  How each part is
pinned (golden vectors, closed forms, library routines, the compiled
reference GSF program) is stated in the docstring of each module.
"""
