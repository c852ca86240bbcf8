// Host-only check of the cluster tables of general_code_amd/cpp/mag.h
// (group_percolating and make_cluster_tables; no GPU call).
// stdin: Lx Ly Lz Lt n, then n records of MAG_CLUSTER_INTS ints.
// stdout: "U length number" per unwrapped length, "W length w0 w1 w2 w3 p"
// per wrapped cluster (p = 1 in the percolating group), "A asymmetry".
#include <cstdio>
#include <iostream>
#include <vector>

#include "mag.h"

using namespace mag_amd;

int main() {
  int d[4];
  long long n = 0;
  if (!(std::cin >> d[0] >> d[1] >> d[2] >> d[3] >> n)) return 2;
  std::vector<int32_t> rec((size_t)n * MAG_CLUSTER_INTS);
  for (auto& v : rec)
    if (!(std::cin >> v)) return 2;
  std::vector<loop_new*> LL = detail::clusters_of(rec.data(), n, d);
  const cluster_tables t = make_cluster_tables(LL, {d[0], d[1], d[2], d[3]});
  free_clusters(LL);
  for (const auto& [len, num] : t.unwrapped) std::cout << "U " << len << " " << num << "\n";
  for (size_t i = 0; i < t.wrapped_lengths.size(); i++) {
    bool p = false;
    for (int k : t.percolating) p |= k == (int)i;
    std::cout << "W " << t.wrapped_lengths[i];
    for (int w : t.wrappings[i]) std::cout << " " << w;
    std::cout << " " << (p ? 1 : 0) << "\n";
  }
  std::printf("A %.17g\n", t.asymmetry);
  return 0;
}
