"""CPU oracle for the filtered-backpropagation path (TEST INFRASTRUCTURE ONLY).

This package is a CPU restatement of ODTbrain's filtered backpropagation
(``odtbrain/_alg3d_bpp.py``, ``_alg3d_bppt.py``, ``_alg2d_bpp.py``,
``util.py``) in numpy/scipy.  It is the checker the GPU implementation is
compared against and the CPU baseline timed beside it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import it; the product package ``odtbrain_amd`` never does.

Substitutions relative to the reference (rounding-level, ~1e-15 relative):
FFTW (pyfftw) -> ``scipy.fft`` with ``workers=num_cores``; numexpr -> numpy.
scipy.ndimage is called exactly as the reference calls it.

Pinning (what checks each function, see tests/test_oracle_*.py):
  * backpropagate_3d      -- reference fixtures alg3d_bpp__test_3d_backprop_phase
                             and _nopadreal (tests/test_alg3d_bpp.py:30-62)
  * mprotate              -- reference fixture alg3d_bpp__test_3d_mprotate
                             (tests/test_alg3d_bpp.py:116-134)
  * backpropagate_2d      -- reference fixtures alg2d_bpp__test_2d_backprop_phase
                             and _full (tests/test_alg2d_bpp.py:13-36)
  * compute_angle_weights_1d -- closed forms of tests/test_util.py:6-32
  * tilted geometry       -- known answers of tests/test_rotation_matrices.py
                             and tests/test_spherecoords_from_angles_and_axis.py
  * backpropagate_3d_tilted -- equality with backpropagate_3d for the y axis
                             (tests/test_alg3d_bppt.py:10-50), which is pinned
                             by the fixtures above
  * bspline (prefilter + interpolation restatement) -- scipy.ndimage on
                             random inputs, plus the mprotate fixture
"""
