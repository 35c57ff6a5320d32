"""CPU oracle for the moljax stiff-PDE hot path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy/SciPy float64 restatement of the reference moljax
algorithms (each function cites the reference file:line it follows).  It is
the checker the HIP path is compared against and the CPU baseline timed by
bench.py.  Only ``tests/``, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import it; the product package ``moljax_amd`` never
does, and has no CPU fallback.

Pinning (what fixes each function independently of this restatement):
  spectral.fft_cn_*            reference-held golden numbers
                               (benchmarks/results/scipy_comparison.json:21,
                               benchmarks/results/bc_matrix.json:8,28,48) and the
                               closed form G^n u0.
  spectral.laplacian_symbol_2d the periodic 5-point stencil applied to Fourier
                               modes (exact eigenvalues).
  spectral.solve_helmholtz_2d  FD residual: (I - dt D lap_h) u = rhs to 1e-13.
  gray_scott.*                 fixed point (u=1, v=0), pure-diffusion closed
                               form (1 - dt/2 D lam)^-2 per step, reaction
                               formula on scalars.
  jvp.*                        golden AD-JVP norm
                               (benchmarks/results/jvp_vs_fd_sweep.json:3) and
                               central finite differences.
"""
