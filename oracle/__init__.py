"""CPU oracle for the particle-mesh gravity path and its scatter-field surrogate.

TEST INFRASTRUCTURE ONLY.  Nothing in ``augur_amd`` imports this package: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or, for ``cpu_baseline``, as the
CPU restatement being timed beside the HIP path).  The product path fails
loudly when its HIP library is missing; it never falls back to this code.

Modules
-------
``pm``        numpy restatement (FP64 unless stated) of the reference algorithm:
              cell keys, stable sort-by-cell, CIC deposit/gather, the 2D
              open-boundary ``pm_acc``, the 3D periodic Green's-function Poisson
              solve with spectral gradient, KDK leapfrog, force-error metrics.
``direct``    FP64 softened direct sum (isolated, the reference's exact_acc),
              FP64 Ewald sum (periodic), energies and momenta.
``torch_cpu`` PyTorch-CPU FP32 port of the plain PM step (the CPU baseline,
              3D periodic and the 2D reference-shape pm_acc) and the FP64
              conv3d / LocalNet3D reference used for the surrogate network.
``scatter2d`` FP64 restatement of the reference's 2D ScatterField step (UNet
              and LocalNet grid nets, learned-kernel FFT with zero and ext
              padding, FD gradient, CIC and TSC stencils, cell-list kNN pair
              term, momfix, Verlet and plain updates) and the reader of its
              shipped weights manifest.

Each function cites the reference file:line it restates.  How each one is
pinned (golden fixture, closed form, brute force) is listed in the test files
``tests/test_oracle_*.py``; functions with no such pin say "parity unpinned"
in their docstring.
"""
