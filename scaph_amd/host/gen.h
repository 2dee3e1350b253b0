// gen.h — host-side synthetic graph generators (the reference ships none;
// its graphs come from an external toolchain).  R-MAT follows the edge
// stream of csrc/rmat_spec.h bit for bit, so a graph generated here equals
// the one the device builder produces from the same parameters.
#ifndef SCAPH_HOST_GEN_H
#define SCAPH_HOST_GEN_H

#include <cstdint>
#include <string>
#include <vector>

#include "graph_io.h"

namespace scaph {

struct RmatParams {
  uint32_t scale = 20;
  uint32_t edge_factor = 16;
  double a = 0.57, b = 0.19, c = 0.19;  // d = 1 - a - b - c
  uint64_t seed = 1;
};

// Power-law out-degree of vertex v: min(cap, floor(x_min * (1 - u)^(-1 /
// (alpha - 1)))) with u the top 53 bits of SplitMix64 word v of seed over
// 2^53 (a floored, capped continuous Pareto with density ~ x^-alpha).
uint32_t zipf_degree(uint64_t seed, uint32_t v, double x_min, double alpha, uint32_t cap);

// u32 quadrant thresholds floor(x * 2^32) (clamped to 2^32 - 1).
void rmat_thresholds(const RmatParams& p, uint32_t* thr_a, uint32_t* thr_ab, uint32_t* thr_abc);

// Directed R-MAT: E = edge_factor * 2^scale raw edges, vertex IDs permuted,
// self-loops and duplicates removed, neighbour lists sorted ascending.
HostCsr generate_rmat(const RmatParams& p, int threads = 0);

// Directed Erdos-Renyi G(n, m): m raw edges with endpoints uniform on [0,n)
// (Lemire's multiply-high mapping of SplitMix64 words), then dedup.
HostCsr generate_er(uint32_t n, uint64_t m, uint64_t seed, int threads = 0);

// Raw sorted-unique edge keys (u<<32|v) -> CSR on n vertices.
HostCsr csr_from_sorted_keys(const std::vector<uint64_t>& keys, uint32_t n);

// Edge weights for the SSSP app: weight of CSR edge k = 1 + word(seed ^
// kWeightSalt, k) mod wmax (uniform on [1, wmax]).
void fill_uniform_weights(HostCsr* g, uint32_t wmax, uint64_t seed);

// Parses "rmat:scale=20,ef=16,a=0.57,b=0.19,c=0.19,seed=1" or
// "er:n=1000,deg=8,seed=3"; either takes "wmax=K" for uniform edge weights
// on [1, K] (fill_uniform_weights).  Returns false on a malformed spec.
bool generate_from_spec(const std::string& spec, HostCsr* out, std::string* err);

// Source selection used by the benchmarks: highest out-degree (lowest id
// on ties).
uint32_t max_degree_vertex(const HostCsr& g);

}  // namespace scaph

#endif
