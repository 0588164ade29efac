"""Algorithmic operation counts and MI355X peaks used for roofline fractions.

The hot path is vector-ALU bound (dependent FP chains, transcendentals), not
HBM bound and not a dense contraction.  Counts are per work unit and frozen
here so the builder and any reviewer compute the same fraction:

* a transcendental (sin, cos, exp, log, sqrt, a division) counts as
  ``TRANSCENDENTAL_FP32 = 20`` FP32 operations (``TRANSCENDENTAL_FP64 = 40``);
* an FMA counts as 2 operations.

Per-unit counts (estimates of the reference algorithm, not of these kernels'
instruction stream):

* sinc^2 cell, one (line, bin): phase ``a (E - E_r)`` (2), reduction to
  ``[-pi/2, pi/2]`` (2), square and scale (4): 8 FP ops, plus one sin and
  one division -> ``8 + 2*20 = 48`` FP32-equivalent operations;
* brem cell, one (segment, bin, element): about 25 FP ops plus log, exp, sqrt,
  exp and expm1 of the escape mean -> ``25 + 5*20 = 125``;
* transport step (LUT, screened Rutherford, midpoint rule, one element, Urban
  straggling): about 90 FP64 ops and 6 transcendentals (3 log, sqrt, sin, cos)
  -> ``90 + 6*40 = 330`` FP64-equivalent operations, plus the Urban draws;
* line prologue, one (segment, reflection, direction): about 220 FP64 ops and
  8 transcendentals -> ``220 + 8*40 = 540`` FP64-equivalent operations;
* coherent field cell, one (line, bin) of the formation-factor route
  (``coherent_stream_jit_kernel.py:153-187``): formation argument (3), the
  factor's ``w^2`` and real/imaginary parts (12), propagation phase (4), the
  phase rotation (6) and two complex multiply-adds (16), plus sin and cos of
  both arguments and one division -> ``41 + 5*20 = 141``.
"""

TRANSCENDENTAL_FP32 = 20
TRANSCENDENTAL_FP64 = 40

SINC_CELL_OPS = 8 + 2 * TRANSCENDENTAL_FP32
BREM_CELL_OPS = 25 + 5 * TRANSCENDENTAL_FP32
TRANSPORT_STEP_OPS_FP64 = 90 + 6 * TRANSCENDENTAL_FP64
LINE_PROLOGUE_OPS_FP64 = 220 + 8 * TRANSCENDENTAL_FP64
COHERENT_CELL_OPS = 41 + 5 * TRANSCENDENTAL_FP32

# MI355X (gfx950) published peaks.
PEAK_FP32_TFLOPS = 157.3  # vector FP32 (also the dense FP32 matrix figure)
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0  # dense MFMA
PEAK_HBM_GBS = 8000.0
