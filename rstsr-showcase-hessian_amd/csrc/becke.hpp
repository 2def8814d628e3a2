// Becke partition weights and their nuclear derivatives on the device (reference:
// src/numint_matmul/becke_partition.rs).  Each grid point belongs to its generating atom and
// moves with it; w = wquad P_g / Z with Becke's cell functions (J. Chem. Phys. 88, 2547 (1988)),
// dw and ddw are total derivatives with respect to the atom positions.
#pragma once
#include <vector>

#include "gemm.hpp"

namespace hh {

class Becke {
 public:
  // BeckeMolTables::new (becke_partition.rs:163-205): atm_coords host [natm, 3], adjustment host
  // [natm, natm] row-major (entry (A, B) is a_AB, antisymmetric).
  Becke(int64_t natm, const double* atm_coords, const double* adjustment, hipStream_t s);

  struct Out {
    double* w = nullptr;    // [ng]
    double* dw = nullptr;   // [natm, 3, ng]
    double* ddw = nullptr;  // [natm, 3, natm, 3, ng]
    const double* contract_w = nullptr;   // [nset_w, ng] -> c [nset_w]
    int64_t nset_w = 0;
    double* c = nullptr;
    const double* contract_dw = nullptr;  // [nset_dw, ng] -> dc [natm, 3, nset_dw]
    int64_t nset_dw = 0;
    double* dc = nullptr;
    const double* contract_ddw = nullptr; // [nset_ddw, ng] -> ddc [natm, 3, natm, 3, nset_ddw]
    int64_t nset_ddw = 0;
    double* ddc = nullptr;
  };
  // becke_partition_with_tables (becke_partition.rs:259-340).  coords device [ng, 3] (row-major
  // triples), wquad device [ng].  by_atom: idx host [natm + 1] cumulative grid boundaries; else
  // idx host [ng] per-grid generating atoms (values >= natm are padding and carry no weight).
  // Outputs and contraction arrays live on the device; null outputs are not produced.
  void partition(const double* coords, int64_t ng, bool by_atom, const int64_t* idx, const double* wquad,
                 int hardness, int deriv, const Out& o);

  int64_t natm() const { return natm_; }
  hipStream_t stream() const { return stream_; }

 private:
  int64_t natm_;
  hipStream_t stream_;
  DevBuf tab_, meta_, scr_, part_, ws_;
};

}  // namespace hh
