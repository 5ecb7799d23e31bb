"""CPU oracle for the effective-sound-speed normal-mode path -- TEST INFRASTRUCTURE ONLY.

This package restates, in FP64 numpy/scipy, what ncpaprop's ESSModeSolver path computes
(profile reader, Steffen splines, c_eff, Sutherland-Bass attenuation, the finite-difference
Helmholtz tridiagonal with its quirks, eigenpairs, perturbative attenuation, 1-D / N x 2-D /
2-D transmission loss, dispersion records and broadband pulse synthesis).  Every function
cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker / the timed CPU baseline.  The product path
(``ncpaprop-release_amd``) never imports or calls anything here.

Pinning (what checks each part; see tests/test_oracle_*.py):
  * profile reader / units          -- round trips on the reference sample profiles, exact
                                       conversion factors.
  * Steffen spline                  -- monotonicity, exact reproduction of node values,
                                       C1 continuity, linear/quadratic data reproduced, and
                                       end-to-end through the reference goldens below.
  * c0 / wind / c_eff / attenuation -- closed forms (c0 = sqrt(1.4*287*T) for ideal-gas
                                       profiles), and through the goldens.
  * matrix + eigenpairs             -- closed form for constant c_eff (ghost-node BC),
                                       dense numpy.linalg.eigh brute force on small Nz,
                                       Sturm counts, and the reference dispersion goldens
                                       (Re k to ~1e-9, psi(zs)psi(zr)).
  * perturbation                    -- closed form sqrt(k^2 + 2 i a) for constant weights.
  * TL 1-D / 2-D                    -- single-mode closed forms, 2-D row at z_r == 1-D,
                                       and the reference tloss_1d goldens (x1000 scaling).
  * broadband synthesis             -- numpy.fft == FFTW_FORWARD, closed-form pulses, the
                                       reference source_spectrum golden.
  * WKB ground-to-ground cut-off    -- hand-worked 20-row grids (stated k_max), the
                                       closed-form tunnelling integral of a linear k_eff^2,
                                       both bypass branches (test_oracle_wkb.py).
  * group speeds, modal starter,    -- constant-c closed form and d omega/dk of the discrete
    attenuation file                   dispersion relation; the lossless field at the source
                                       range; Steffen on the end-extended table
                                       (test_oracle_options.py).
  * C restatement (ess_port.c)      -- the same path in plain C (the timed CPU baseline);
                                       checked against this package's functions on the same
                                       inputs (test_oracle_port.py).
  * wide-angle modes (wmod.py)      -- the reference WMod goldens wtloss_1d{,.lossless}.nm,
                                       the reference's 2N x 2N linearisation solved densely,
                                       the constant-c / constant-Mach closed form, and the
                                       zero-wind reduction to the ESS operator.
"""
