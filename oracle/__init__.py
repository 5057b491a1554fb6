"""CPU oracle for the PyHermes field pipeline -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference algorithms of the field
pipeline (projection, window kernels, FFT convolution, pair products, the
spherical-harmonic 3PCF contraction and the Monte-Carlo triplet estimator).
Every function cites the reference file:line it follows
(paths relative to the PyHermes source tree).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product package (``pyhermes_amd``) never imports it and has no
CPU fallback.

Parity pinning (what fixes each function independently of retyping it):

* ``scaling.sample_scaling_function``: invariants (length 3072, phi[0] = 0,
  sum = 1024, partition of unity) and the closed-form DTFT product of the
  cascade, which also pins ``scaling.fourier_power_spectrum``.  The db2 filter
  literals are the ones PyWavelets prints for ``Wavelet('db2').rec_lo``; only
  an import of PyWavelets (absent here) pins phi bit for bit, so phi is pinned
  to ~1e-14 by the invariants, not bit-exactly.
* ``scaling.cal_gamma``: partition of unity (sum of Gamma = 1) and a brute
  force evaluation with explicit modular indexing.
* ``projection``: brute force ``np.add.at`` on tiny grids, sum(eps) = sum(w),
  whole-cell shift equivariance, closed-form cell/fraction indices.
* ``special``: scipy.special.spherical_jn / sph_harm_y (the reference's own
  validation route, legendre_windows.py:18-38).
* ``windows``: closed forms at k = 0, octant fold == explicit 8-alias rFFT
  builder, multipole conjugation symmetry.
* ``conv`` / ``pair``: scipy.fft and the Parseval identity; real-space brute
  force pair sums on tiny grids.
* ``multipole``: brute-force 27x27 triple sum on tiny grids and the constant
  field closed form v = 4*pi*a*b*c.
* ``mt19937`` (the rotation draws of ``mc``): Matsumoto & Nishimura's
  init_genrand + genrand_res53, the generator Numba's np.random.seed/rand
  implement, pinned by the C++ standard's known value (std::mt19937 seeded 5489
  gives 4123659995 as its 10,000th output) and by libstdc++'s std::mt19937
  for other seeds; everything downstream of the draws is pinned by brute force.
* ``multipole.xi_multipole``: the Gauss-Legendre mu-integral of the ring-window
  (s, mu) pair products, an independent route to the xi_l prefactor.
"""
