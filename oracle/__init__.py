"""CPU restatement of the reference algorithms for the preconditioner + PCG path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker (or as the timed
CPU baseline).  The product path (mlpcg_amd) never imports it and fails loudly when
its HIP library is missing.

Modules
-------
model     FP64 torch restatement of sm_model_3d_py.py / sm_model_3d.py (SM block,
          transpose, linear, level transfers, SmallSMModelDn3D and the SPD variants)
          plus pure-Python brute-force forms for tiny grids.
solvers   npcg / npcg_flex / npsd (cg_tests.py:26-284) and test.py:42-51 predict.
laplacian loop-form restatement of lib/discrete_laplacian.py:122-160 for small grids.
train     residual loss of loss_functions.py:18-23 and the configured squared form.

Parity pins (each function's check lives in tests/test_oracle_*.py):
  sm_block / sm_block_trans / stencil_weights : brute force on tiny grids, F.conv3d
      identity, special cases (identity stencil, constant image -> conv3d), adjoint.
  sm_linear : brute-force mean patch, F.conv3d(...).mean() closed form.
  pool / upsample : F.avg_pool3d / F.interpolate (library) + hand-computed 1-D cases.
  model : composition checked against a brute-force composition on 8^3, linearity.
  npcg  : P = I reproduces scipy.sparse.linalg.cg residual history; P = A^-1 -> 1 step.
  laplacian : hand-computed 3^3 / 4^3 cases, symmetry, row sums = #air neighbours.
"""
