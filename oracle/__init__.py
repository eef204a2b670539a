"""Oracle: a CPU float64 restatement of tf-pwa's unbinned-likelihood hot path.

TEST INFRASTRUCTURE. This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import it. The
product path (``tf-pwa_amd/``) never imports, calls or links anything under ``oracle/``.

The restatement is written with eager torch CPU tensor ops in float64/complex128, one op
for each TensorFlow op of the reference, and ``torch.autograd`` plays the role of
``tf.GradientTape``. See ``oracle/tfpwa_oracle.py`` for the per-function citations and the
parity status of each piece.
"""
