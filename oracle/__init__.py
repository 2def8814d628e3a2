"""FP64 CPU oracle for the hesship hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its C ABI,
or the `hesship` host package) imports, links or calls this package.  Only
`tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
may use it, and only as the checker.

Each function restates the reference algorithm (ajz34/rstsr-showcase-hessian,
Rust, column-major tensors) in NumPy FP64 and cites the reference file:line it
follows.  Tensors here are indexed with the reference's logical index order
(e.g. ``bra[mu, i, a]``); memory order is irrelevant inside the oracle.

Pinning (what checks each oracle function independently of its own formula):

* ``rijk.pack_triu_tilde`` / ``unpack_triu``: round trip and the explicit
  diagonal/off-diagonal rule, tests/test_oracle_rijk.py.
* ``rijk.response_bra_separated`` (optimized form): equals the brute-force
  naive einsum form (``rijk.response_bra_naive``, built from full j3c and an
  explicit j2c inverse) for RHF and UHF, J-only/K-only/full, which is the
  reference's own optimized==naive invariant; plus linearity and the
  J-depends-only-on-symmetric-dm1 property.
* ``rijk.cderi_from_j3c``: ``B B^T == j3c j2c^-1 j3c^T``.
* ``krylov.krylov_block``: direct solve on the reference's closed-form test
  system (restart and non-convergence cases included).
* ``cphf``: CPHF residual closed form ``U + resp(U)/de == rhs``, and
  finalize/get_cphf_hess against an explicit loop on tiny sizes.
* ``numint``: brute-force per-grid-point loops on tiny grids, and the
  full-potential contraction ``F[mu,nu] @ bra`` equal to the bra-transformed
  form.
* ``xceff``: chain rule checked by torch.autograd (FP64, CPU) of an analytic
  GGA/meta-GGA energy density.

No reference golden vector for this path survives in the reference tree
(every ``.npz`` input is absent), so no function is pinned by a reference
number; each is pinned by the mathematics as listed.
"""
