// host_tgtp.hpp -- host helpers of the tgtp pipeline (host_tgtp.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace tgtp_host {

// get_tau of aesc.h:129-135 with glibc log, bit-exact with the reference.
int tau_bound(uint64_t deg_u, uint64_t deg_v, float eps, double lam, float del, float ups);

// tau of the undirected edge (u, v), aesc.h:182-204.
int edge_tau(const uint32_t* idx, const float* X, int omega, const double* lam, uint64_t m,
             float eps, uint32_t u, uint32_t v);

// [k_rows][omega] table of lambda_i^(2(k+1)) by repeated products.
std::vector<double> power_table(const double* lam, int omega, uint32_t k_rows);

// For each CSR position t, the position whose pred lands in rev_preds[t]
// during symmetrization (-1 if none).  The reference's loader always stores
// sorted rows holding both directions of every edge; a caller's CSR may
// not, and then its lower_bound would find a wrong position or one past v's
// row.  Returns 0, 1 for an unsorted row, 2 for an edge (u, v) without
// (v, u).
int symmetrize_sources(const uint32_t* idx, const uint32_t* adj, uint64_t n, uint64_t m,
                       std::vector<int32_t>& src);

}  // namespace tgtp_host
