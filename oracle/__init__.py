"""FP64 CPU oracle for the compact-binary likelihood hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package (``bilby_amd``)
imports, links or executes anything under ``oracle/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
use it, and only as the checker / the timed CPU baseline.

The oracle is a plain numpy (float64 / complex128) restatement of bilby's
per-sample dataflow, one parameter dict in, one float out.  Each function
cites the reference file:line it follows (paths relative to the bilby source
tree).  scipy is used only where bilby itself calls scipy
(``special.i0e``, ``special.logsumexp``, ``RectBivariateSpline``).

Pinning status (what checks each function, see tests/test_oracle_*.py):

* series / noise / PSD interpolation / masked inner products: pinned by the
  reference's known-answer values (noise lnL -4014.1787704539474,
  inner product 239.87768033598326, matched-filter SNR 25.510869054168282).
* GMST / antenna pattern / time delay: pinned by closed-form invariants
  (F+^2+Fx^2 independent of psi, |dt| <= R_earth/c, GMST rate) and by
  sidereal-time known values.
* sine-Gaussian: pinned by its closed-form <h|h> for a flat PSD.
* TaylorF2: partially pinned (Newtonian SPA amplitude/phase vs an FFT of a
  numerically generated Newtonian chirp; every non-spinning phase
  coefficient, 1PN-3.5PN with the 2.5PN and 3PN logarithms, vs the 3PN
  energy / 3.5PN flux re-expansion; the spin-orbit terms at 1.5-3.5PN vs
  the spin-orbit energy and flux).  The 2PN spin-spin terms are
  transcribed only.  Parity with lalsimulation is
  UNPINNED (lalsuite is not available).
* time marginalisation: pinned by brute-force DFT and by brute-force
  trapezoid integration of the un-marginalised likelihood.
* phase marginalisation: pinned by numerical integration over phase.
* distance marginalisation: pinned by direct 10^4-point integration.
* ROQ: pinned by brute-force evaluation of the weights' defining sums.
* relative binning: pinned at the fiducial point (equal to the full
  likelihood), by the piecewise-linear reconstruction summed on the grid,
  and by the reference's draw property (tests/test_oracle_relative.py).
"""
