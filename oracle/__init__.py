"""CPU oracle for the probabilistic-circuit hot path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this package, and only as the checker / CPU baseline.  The product path
(``pyjuice_amd``) never imports it.

Contents
--------
* ``dag_eval``: FP64 numpy evaluation directly on the node DAG (no compiled indices):
  forward log-likelihoods, flows, parameter flows, mini-batch EM and ancestral sampling
  from caller-supplied uniforms.  Follows the maths of the reference's layers
  (``pyjuice/layer/sum_layer.py:454-568``, ``:1569-1730``, ``:2049-2169``;
  ``pyjuice/layer/prod_layer.py:398-466``; ``pyjuice/nodes/distributions/categorical.py:83-245``,
  ``bernoulli.py:65-101``; ``pyjuice/model/backend/par_update.py:173-256``;
  ``pyjuice/queries/sample.py:38-140``).
* ``pc_cpu.c`` + ``pc_cpu.py``: C / OpenMP restatement of forward, flow backward and EM on the
  compiled index tables, built in FP64 (an oracle fast enough for full-size circuits, pinned
  to ``dag_eval`` by ``tests/test_oracle_cpu.py``) and in FP32 (the CPU baseline of
  ``bench.py``) by ``build_cpu.py``.

Pinning (see ``tests/test_oracle.py``): every function is checked against what the
mathematics fixes -- exhaustive enumeration (sum of probabilities = 1), central finite
differences of the log-likelihood (parameter flows = theta * dLL/dtheta), flow
conservation, the closed-form EM fixed point.  The compiler is pinned separately against the
literal index tables transcribed from the reference's tests (``tests/golden/``).
"""
