// mag.h -- C++ drop-in for the MAG call surface of ilyakudrov/general_code
// (lib/cpu/include/mag.h), running on the MI355X through the C ABI of
// include/mag_gpu.h.
//
// The functions are templates over the lattice and spin types, so an app
// can pass either the reference's own objects
//   Data::LatticeData<DataPatternLexicographical, su2>   (data.h:78-237)
//   std::vector<spin>                                     (matrix.h:773-787)
// or the layout-identical mirror types below; only the memory layout is
// used: su2 = 4 doubles {a0,a1,a2,a3}, spin = 3 doubles, links in the
// order 4*site + mu, site = x + Lx*(y + Ly*(z + Lz*t)) (indexing.h:71-126).
// The reference's su3 is an Eigen::Matrix3cd (column-major); the mirror su3
// below is row-major (the file order of matrix.h:236-254); both are
// accepted (see su3_row_major).
//
// Differences from the reference, by design:
//   * sweeps run in checkerboard order on the GPU (the reference's in-place
//     lexicographic OpenMP loops race, mag.cpp:1386, 1491);
//   * random numbers come from a Philox stream keyed by an explicit seed
//     (the reference seeds with time(NULL), mag.cpp:145, 764);
//   * failures throw mag_amd::error instead of printing and continuing;
//   * the maximisation loops have an iteration cap (max_iter).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <tuple>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mag_gpu.h"

namespace mag_amd {

struct error : std::runtime_error {
  int code;
  error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != MAG_OK) throw error(rc, mag_last_error());
}

// ---------------------------------------------------------------- value types
// su2: a0 + i (a1 s1 + a2 s2 + a3 s3) (matrix.h:8-13)
struct su2 {
  static constexpr int data_size = 4;
  double a0 = 1, a1 = 0, a2 = 0, a3 = 0;
};
// spin (matrix.h:773-787); the default is (0, 0, 1) as in the reference
struct spin {
  double a1 = 0, a2 = 0, a3 = 1;
};
// su3, row-major (re, im) (matrix.h:236-254)
struct su3 {
  static constexpr int data_size = 18;
  double v[18] = {1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0};
};
// abelian link r e^{i phi} (matrix.h:119-130) and the three U(1) angles of
// an su3_angles link (matrix.h:579-585), the layouts calculate_current reads
struct abelian {
  double r = 1, phi = 0;
};
struct su3_angles {
  std::array<double, 3> matrix{0, 0, 0};
};

// link index 4*site + mu (indexing.h:71-126)
class DataPatternLexicographical {
 public:
  std::array<int, 4> lat_dim;
  explicit DataPatternLexicographical(const std::array<int, 4>& d) : lat_dim(d) {}
  int get_lattice_size() const { return lat_dim[0] * lat_dim[1] * lat_dim[2] * lat_dim[3]; }
  int get_data_size() const { return get_lattice_size() * 4; }
};

namespace Data {
// data.h:78-237 (storage part)
template <class DataPattern, class MatrixType>
class LatticeData {
 public:
  typedef MatrixType matrix_type;
  typedef DataPattern data_pattern_type;
  std::vector<MatrixType> array;
  std::array<int, 4> lat_dim;
  explicit LatticeData(const std::array<int, 4>& d)
      : array((size_t)d[0] * d[1] * d[2] * d[3] * 4), lat_dim(d) {}
  const MatrixType& operator[](int i) const { return array[i]; }
  MatrixType& operator[](int i) { return array[i]; }
};
}  // namespace Data

// su3 storage order of a matrix type: the mirror su3 is row-major; any
// other 18-double type is taken to be Eigen::Matrix3cd (column-major).
template <class M>
struct su3_row_major : std::is_same<M, su3> {};

namespace detail {

template <class Lattice>
const double* su2_links(const Lattice& conf) {
  using M = typename Lattice::matrix_type;
  static_assert(sizeof(M) == 4 * sizeof(double), "su2 link must be 4 doubles");
  return reinterpret_cast<const double*>(conf.array.data());
}
template <class Lattice>
double* su2_links(Lattice& conf) {
  using M = typename Lattice::matrix_type;
  static_assert(sizeof(M) == 4 * sizeof(double), "su2 link must be 4 doubles");
  return reinterpret_cast<double*>(conf.array.data());
}
template <class S>
double* spin_data(std::vector<S>& s) {
  static_assert(sizeof(S) == 3 * sizeof(double), "spin must be 3 doubles");
  return reinterpret_cast<double*>(s.data());
}
template <class Lattice>
void dims_of(const Lattice& conf, int* d) {
  for (int i = 0; i < 4; i++) d[i] = conf.lat_dim[i];
}
// su3 links in the ABI's row-major order
template <class Lattice>
std::vector<double> su3_links(const Lattice& conf) {
  using M = typename Lattice::matrix_type;
  static_assert(sizeof(M) == 18 * sizeof(double), "su3 link must be 18 doubles");
  std::vector<double> out(conf.array.size() * 18);
  const double* in = reinterpret_cast<const double*>(conf.array.data());
  for (size_t l = 0; l < conf.array.size(); l++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        const size_t src = su3_row_major<M>::value ? (size_t)(3 * i + j) : (size_t)(i + 3 * j);
        out[18 * l + 2 * (3 * i + j)] = in[18 * l + 2 * src];
        out[18 * l + 2 * (3 * i + j) + 1] = in[18 * l + 2 * src + 1];
      }
  return out;
}
template <class Lattice>
void su3_store(Lattice& conf, const std::vector<double>& rm) {
  using M = typename Lattice::matrix_type;
  double* o = reinterpret_cast<double*>(conf.array.data());
  for (size_t l = 0; l < conf.array.size(); l++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        const size_t dst = su3_row_major<M>::value ? (size_t)(3 * i + j) : (size_t)(i + 3 * j);
        o[18 * l + 2 * dst] = rm[18 * l + 2 * (3 * i + j)];
        o[18 * l + 2 * dst + 1] = rm[18 * l + 2 * (3 * i + j) + 1];
      }
}

// A spin session for one call: links (and spins) uploaded, freed on exit.
class Session {
 public:
  template <class Lattice>
  explicit Session(const Lattice& conf) {
    int d[4];
    dims_of(conf, d);
    int dev = 0;
    check(mag_device_count(&dev));
    check(mag_spin_session_create(d, 0, &s_));
    int rc = mag_spin_session_upload_links(s_, su2_links(conf));
    if (rc != MAG_OK) {
      mag_spin_session_destroy(s_);
      throw error(rc, mag_last_error());
    }
  }
  ~Session() { mag_spin_session_destroy(s_); }
  Session(const Session&) = delete;
  Session& operator=(const Session&) = delete;
  mag_spin_session* get() const { return s_; }

 private:
  mag_spin_session* s_ = nullptr;
};

}  // namespace detail

// ---------------------------------------------------------------- mag.h surface
// generate_spins_uniform (mag.h:179-199)
template <class SpinT = spin, class DataPattern>
std::vector<SpinT> generate_spins_uniform(DataPattern& pattern, uint64_t seed) {
  int d[4];
  for (int i = 0; i < 4; i++) d[i] = pattern.lat_dim[i];
  std::vector<SpinT> s((size_t)d[0] * d[1] * d[2] * d[3]);
  check(mag_su2_generate_spins_uniform(d, seed, detail::spin_data(s)));
  return s;
}

// read_spins (mag.h:12-22, mag.cpp:54-88): V spins of 3 raw doubles each,
// lexicographic site order.  The reference prints an error and goes on for
// a short file; here that is an error.  (Its overloads without a pattern
// size the read by the process globals x_size.., which apps/mag_su2 never
// sets, so they read nothing; this one always reads the lattice's V spins.)
template <class SpinT = spin, class DataPattern>
std::vector<SpinT> read_spins(const std::string& path, const DataPattern& pattern) {
  size_t V = 1;
  for (int i = 0; i < 4; i++) V *= (size_t)pattern.lat_dim[i];
  std::vector<SpinT> s(V);
  std::ifstream f(path, std::ios::binary);
  if (!f || !f.read(reinterpret_cast<char*>(detail::spin_data(s)), (std::streamsize)(3 * V * sizeof(double))))
    throw error(MAG_ERR_ARG, "read_spins: cannot read " + std::to_string(3 * V) + " doubles from " + path);
  return s;
}

// write_spins (mag.h:12-22, mag.cpp:107-141): 3 raw doubles per spin
template <class SpinT>
void write_spins(const std::string& path, std::vector<SpinT>& spins) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(detail::spin_data(spins)),
          (std::streamsize)(3 * spins.size() * sizeof(double)));
  if (!f) throw error(MAG_ERR_ARG, "write_spins: cannot write " + path);
}

// MAG_functional_su2 (mag.h:69-70, mag.cpp:993-1001)
template <class Lattice>
double MAG_functional_su2(const Lattice& conf) {
  int d[4];
  detail::dims_of(conf, d);
  double F;
  check(mag_su2_functional(d, detail::su2_links(conf), &F));
  return F;
}

// MAG_functional_su2_spin (mag.h:58-60, mag.cpp:925-953)
template <class Lattice, class SpinT>
double MAG_functional_su2_spin(const Lattice& conf, std::vector<SpinT>& spins) {
  int d[4];
  detail::dims_of(conf, d);
  double F;
  check(mag_su2_functional_spin(d, detail::su2_links(conf), detail::spin_data(spins), &F));
  return F;
}

// gauge_tranformation_spins (mag.h:80-82, mag.cpp:1045-1068)
template <class Lattice, class SpinT>
void gauge_tranformation_spins(Lattice& conf, std::vector<SpinT>& spins) {
  int d[4];
  detail::dims_of(conf, d);
  check(mag_su2_gauge_transform_spins(d, detail::su2_links(conf), detail::spin_data(spins)));
}

// make_simulated_annealing (mag.h:119-122, mag.cpp:1585-1603)
template <class Lattice, class SpinT>
void make_simulated_annealing(const Lattice& conf, std::vector<SpinT>& spins, double T_init,
                              double T_final, double T_step, int OR_steps,
                              int thermalization_steps, uint64_t seed = 1) {
  detail::Session s(conf);
  check(mag_spin_session_set_spins(s.get(), detail::spin_data(spins)));
  check(mag_spin_session_simulated_annealing(s.get(), T_init, T_final, T_step, OR_steps,
                                             thermalization_steps, seed));
  check(mag_spin_session_get_spins(s.get(), detail::spin_data(spins)));
}

// make_maximization_final (mag.h:166-169, mag.cpp:1800-1816)
template <class Lattice, class SpinT>
int make_maximization_final(const Lattice& conf, std::vector<SpinT>& spins, int OR_steps,
                            double tolerance_maximal, double tolerance_average,
                            int max_iter = 1000000) {
  detail::Session s(conf);
  check(mag_spin_session_set_spins(s.get(), detail::spin_data(spins)));
  int it = 0;
  const int rc = mag_spin_session_maximize_final(s.get(), OR_steps, tolerance_maximal,
                                                 tolerance_average, max_iter, &it);
  check(mag_spin_session_get_spins(s.get(), detail::spin_data(spins)));
  check(rc);
  return it;
}

// make_maximization_approximate (mag.h:151-153, mag.cpp:1733-1750)
template <class Lattice, class SpinT>
int make_maximization_approximate(const Lattice& conf, std::vector<SpinT>& spins, int OR_steps,
                                  int tolerance_digits, int max_iter = 1000000) {
  detail::Session s(conf);
  check(mag_spin_session_set_spins(s.get(), detail::spin_data(spins)));
  int it = 0;
  const int rc = mag_spin_session_maximize_approximate(s.get(), OR_steps, tolerance_digits,
                                                       max_iter, &it);
  check(mag_spin_session_get_spins(s.get(), detail::spin_data(spins)));
  check(rc);
  return it;
}

// mag_functional_su3 (mag.h:177, mag.cpp:1837-1862)
template <class Lattice>
double mag_functional_su3(const Lattice& conf) {
  int d[4];
  detail::dims_of(conf, d);
  const std::vector<double> rm = detail::su3_links(conf);
  double F;
  check(mag_su3_functional(d, rm.data(), &F));
  return F;
}

// plaket (plaket.h:84-101) for su2 (4 doubles) or su3 (18 doubles) links
template <class Lattice>
double plaket(const Lattice& conf) {
  int d[4];
  detail::dims_of(conf, d);
  double P;
  if constexpr (sizeof(typename Lattice::matrix_type) == 4 * sizeof(double)) {
    check(mag_plaquette(MAG_SU2, d, detail::su2_links(conf), &P));
  } else {
    const std::vector<double> rm = detail::su3_links(conf);
    check(mag_plaquette(MAG_SU3, d, rm.data(), &P));
  }
  return P;
}

// ---------------------------------------------------------------- link form
// The link-form path: checkerboard omega-overrelaxation on the links
// (SU(3): Cabibbo-Marinari), FP64 or mixed precision, in place.  Returns
// the number of sweeps; the functional history is in *history if given.
// With `currents`, also the monopole currents of the fixed links'
// abelian projection, computed before the links leave the device
// (calculate_current's layout: J[4*site+mu], SU(3) J[colour][4*site+mu]).
template <class Lattice>
int fix_mag_links(Lattice& conf, double omega, double tol, int max_sweeps, int precision,
                  std::vector<double>* history = nullptr, std::vector<double>* currents = nullptr) {
  mag_params p;
  mag_params_default(&p);
  constexpr bool is3 = sizeof(typename Lattice::matrix_type) == 18 * sizeof(double);
  p.group = is3 ? MAG_SU3 : MAG_SU2;
  detail::dims_of(conf, p.dims);
  p.omega = omega;
  p.tol = tol;
  p.max_sweeps = max_sweeps;
  p.precision = precision;
  mag_job* job = nullptr;
  check(mag_job_create(&p, &job));
  struct Guard {
    mag_job* j;
    ~Guard() { mag_job_destroy(j); }
  } guard{job};
  std::vector<double> rm;
  if constexpr (is3) {
    rm = detail::su3_links(conf);
    check(mag_job_upload_links(job, rm.data()));
  } else {
    check(mag_job_upload_links(job, detail::su2_links(conf)));
  }
  const int rc = mag_job_run(job);
  if (rc != MAG_OK && rc != MAG_ERR_NOT_CONVERGED) check(rc);
  int sweeps = 0, conv = 0;
  double F = 0;
  check(mag_job_status(job, &sweeps, &conv, &F));
  if constexpr (is3) {
    check(mag_job_download_links(job, rm.data()));
    detail::su3_store(conf, rm);
  } else {
    check(mag_job_download_links(job, detail::su2_links(conf)));
  }
  if (history) {
    history->assign(sweeps, 0.0);
    int n = 0;
    check(mag_job_history(job, history->data(), nullptr, nullptr, sweeps, &n));
  }
  if (currents) {
    currents->assign(conf.array.size() * (is3 ? 3 : 1), 0.0);
    check(mag_job_monopole_currents(job, 0, currents->data()));
  }
  check(rc);
  return sweeps;
}

// ---------------------------------------------------------------- monopoles
// calculate_current (monopoles.cpp:646-687) on the GPU.  The lattice holds
// abelian links {r, phi} (16 bytes; the reference's
// LatticeData<DataPatternLexicographical, abelian>) or su3_angles links (24
// bytes); returns J[4*site + mu] (abelian) or J[colour][4*site + mu]
// (su3_angles), as the reference does.  Angles must be finite with
// |angle| <= 1000 (the reference's unbounded wrap loops are not kept).
template <class Lattice>
auto calculate_current(const Lattice& conf) {
  using M = typename Lattice::matrix_type;
  int d[4];
  detail::dims_of(conf, d);
  const size_t n = conf.array.size();
  const double* in = reinterpret_cast<const double*>(conf.array.data());
  if constexpr (sizeof(M) == 2 * sizeof(double)) {
    std::vector<double> phi(n), J(n);
    for (size_t l = 0; l < n; l++) phi[l] = in[2 * l + 1];
    check(mag_u1_monopole_currents(d, phi.data(), J.data()));
    return J;
  } else {
    static_assert(sizeof(M) == 3 * sizeof(double), "abelian (r, phi) or su3_angles links");
    std::vector<double> J(3 * n);
    check(mag_su3_monopole_currents(d, in, J.data()));
    std::vector<std::vector<double>> out(3);
    for (int c = 0; c < 3; c++) out[c].assign(J.begin() + c * n, J.begin() + (c + 1) * n);
    return out;
  }
}

// ---------------------------------------------------------------- clusters
// A monopole cluster as the apps use it.  The reference's cluster is the
// root of a tree of loop_new nodes (loop.h:21-32) that the apps only walk
// through cluster_length, length_mu and currents_directions
// (monopoles.cpp:1343-1449); the mirror's loop_new carries those sums,
// computed on the GPU, with the coordinate of the cluster's first site.
struct loop_new {
  std::array<int, 4> coordinate{0, 0, 0, 0};
  int length = 0;                          // sum of |charge|
  std::array<int, 4> lengths_mu{0, 0, 0, 0};  // sum of charges per direction
  int space = 0, time = 0;                 // |charge| over mu < 3, mu = 3
  int sites = 0;                           // distinct sites (site_number)
};

namespace detail {
inline std::vector<loop_new*> clusters_of(const int32_t* r, int64_t n, const int* d) {
  std::vector<loop_new*> LL;
  LL.reserve((size_t)n);
  for (int64_t i = 0; i < n; i++, r += MAG_CLUSTER_INTS) {
    auto* c = new loop_new;
    int s = r[0];
    for (int mu = 0; mu < 4; mu++) {
      c->coordinate[mu] = s % d[mu];
      s /= d[mu];
    }
    c->length = r[1];
    for (int mu = 0; mu < 4; mu++) c->lengths_mu[mu] = r[2 + mu];
    c->space = r[6];
    c->time = r[7];
    c->sites = r[8];
    LL.push_back(c);
  }
  return LL;
}
// one cluster entry point, called with the exact capacity after a first
// guess that was short
template <class Call>
std::vector<std::vector<loop_new*>> run_clusters(Call call, int ncol, const int* d) {
  int64_t cap = 1 << 12;
  for (int attempt = 0; attempt < 2; attempt++) {
    std::vector<int32_t> rec((size_t)ncol * cap * MAG_CLUSTER_INTS);
    std::vector<int64_t> n(ncol, 0);
    const int rc = call(rec.data(), cap, n.data());
    int64_t most = 0;
    for (int64_t v : n) most = v > most ? v : most;
    if (rc == MAG_ERR_ARG && most > cap) {
      cap = most;
      continue;
    }
    check(rc);
    std::vector<std::vector<loop_new*>> out(ncol);
    for (int k = 0; k < ncol; k++)
      out[k] = clusters_of(rec.data() + (size_t)k * cap * MAG_CLUSTER_INTS, n[k], d);
    return out;
  }
  throw error(MAG_ERR_ARG, "cluster count changed between calls");
}
}  // namespace detail

// calculate_clusters(J, data_pattern) (monopoles.cpp:1277-1303) on the GPU:
// the clusters of the currents J[4*site + mu] in the reference's order.
// J is left as it is (the reference zeroes it while walking).  The nodes
// belong to the caller, as in the reference; free_clusters deletes them.
template <class T>
std::vector<loop_new*> calculate_clusters(const std::vector<T>& J, const DataPatternLexicographical& dp) {
  const int d[4] = {dp.lat_dim[0], dp.lat_dim[1], dp.lat_dim[2], dp.lat_dim[3]};
  if (J.size() != (size_t)dp.get_data_size())
    throw error(MAG_ERR_ARG, "J has " + std::to_string(J.size()) + " values, the lattice 4V = " +
                                 std::to_string(dp.get_data_size()));
  const std::vector<double> Jd(J.begin(), J.end());
  return detail::run_clusters(
      [&](int32_t* rec, int64_t cap, int64_t* n) { return mag_current_clusters(d, 1, Jd.data(), rec, cap, n); },
      1, d)[0];
}

inline void free_clusters(std::vector<loop_new*>& LL) {
  for (loop_new* c : LL) delete c;
  LL.clear();
}
// the observables of monopoles.cpp:1343-1449 and 1510-1531
inline int cluster_length(const loop_new* c) { return c->length; }
inline std::vector<int> length_mu(const loop_new* c) {
  return {c->lengths_mu[0], c->lengths_mu[1], c->lengths_mu[2], c->lengths_mu[3]};
}
inline std::tuple<int, int> currents_directions(const loop_new* c) { return {c->space, c->time}; }
inline int site_number(const loop_new* c) { return c->sites; }

// group_percolating (monopoles.cpp:1533-1590): the first group of wrapped
// clusters whose winding vectors sum to zero, searched over groups of 2, 3,
// ... in the reference's enumeration.  That enumeration keeps cluster 0 in
// every group; it runs the last position to the end, then moves up by one
// the innermost position i >= 1 whose successor is more than 2 ahead and
// sets the following positions right after it; when there is no such
// position the groups of this size are done.  (So some groups, e.g.
// {0, m-2, m-1} of three out of m, are never tried -- kept, since the app's
// percolating labels depend on it.)
inline std::vector<int> group_percolating(const std::vector<std::vector<int>>& w) {
  const int m = (int)w.size();
  for (int order = 2; order <= m; order++) {
    std::vector<int> pos(order);
    for (int i = 0; i < order; i++) pos[i] = i;
    for (;;) {
      bool zero = true;
      for (int mu = 0; mu < 4 && zero; mu++) {
        int sum = 0;
        for (int i : pos) sum += w[i][mu];
        zero = sum == 0;
      }
      if (zero) return pos;
      if (pos[order - 1] < m - 1) {
        pos[order - 1]++;
        continue;
      }
      int i = order - 2;
      while (i > 0 && pos[i + 1] - pos[i] <= 2) i--;
      if (i <= 0) break;
      pos[i]++;
      for (int k = i + 1; k < order; k++) pos[k] = pos[i] + k - i;
    }
  }
  return {};
}

// The tables apps/monopoles_su2/monopoles_su2.cpp:110-197 (and per colour
// apps/monopoles_su3/monopoles_su3.cpp:117-213) write for one colour's
// clusters: unwrapped clusters counted by length; wrapped clusters (some
// length_mu != 0) with their windings length_mu / L (C++ division), longest
// first -- equal lengths keep the cluster order (the reference's std::sort
// leaves their order unspecified) -- and which of them form the
// percolating group; the asymmetry (space/3 - time) / (space/3 + time).
struct cluster_tables {
  std::map<int, int> unwrapped;
  std::vector<int> wrapped_lengths;
  std::vector<std::vector<int>> wrappings;
  std::vector<int> percolating;  // positions in wrapped_lengths
  double asymmetry = 0;
};
inline cluster_tables make_cluster_tables(const std::vector<loop_new*>& LL, const std::array<int, 4>& dims) {
  cluster_tables t;
  int space = 0, time = 0;
  std::vector<std::pair<int, std::vector<int>>> wrapped;
  for (const loop_new* c : LL) {
    space += c->space;
    time += c->time;
    const std::vector<int> lmu = length_mu(c);
    if (lmu[0] == 0 && lmu[1] == 0 && lmu[2] == 0 && lmu[3] == 0) {
      t.unwrapped[c->length]++;
    } else {
      wrapped.push_back({c->length, {lmu[0] / dims[0], lmu[1] / dims[1], lmu[2] / dims[2], lmu[3] / dims[3]}});
    }
  }
  std::stable_sort(wrapped.begin(), wrapped.end(),
                   [](const auto& a, const auto& b) { return a.first > b.first; });
  for (auto& w : wrapped) {
    t.wrapped_lengths.push_back(w.first);
    t.wrappings.push_back(w.second);
  }
  if (t.wrappings.size() > 1) t.percolating = group_percolating(t.wrappings);
  t.asymmetry = (space / 3. - time) / (space / 3. + time);
  return t;
}

// ---------------------------------------------------------------- file formats
namespace detail {
// LIME records (the container of ILDG files): a 144-byte header -- magic
// 0x456789ab (big-endian 32 bit), version (16 bit), a flag byte with
// message-begin 0x80 / message-end 0x40, the data length (big-endian 64
// bit) at byte 8 and a NUL-padded 128-byte type string at byte 16 --, then
// the data padded to a multiple of 8 bytes.
constexpr uint32_t kLimeMagic = 0x456789abu;
inline uint64_t be_load(const unsigned char* p, int n) {
  uint64_t v = 0;
  for (int i = 0; i < n; i++) v = (v << 8) | p[i];
  return v;
}
inline void be_store(unsigned char* p, uint64_t v, int n) {
  for (int i = n - 1; i >= 0; i--, v >>= 8) p[i] = (unsigned char)(v & 0xff);
}
inline double be_double(const unsigned char* p) {
  const uint64_t u = be_load(p, 8);
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}
// the data of the first record of type `type`, as read
inline std::vector<unsigned char> lime_record(std::ifstream& in, const std::string& path,
                                              const std::string& type) {
  for (;;) {
    unsigned char h[144];
    if (!in.read(reinterpret_cast<char*>(h), sizeof(h)))
      throw error(MAG_ERR_ARG, "no " + type + " record in " + path);
    if (be_load(h, 4) != kLimeMagic) throw error(MAG_ERR_ARG, "not a LIME file: " + path);
    const uint64_t len = be_load(h + 8, 8);
    const std::string t(reinterpret_cast<const char*>(h + 16), strnlen(reinterpret_cast<const char*>(h + 16), 128));
    const uint64_t padded = (len + 7) / 8 * 8;
    if (t == type) {
      std::vector<unsigned char> d(len);
      if (!in.read(reinterpret_cast<char*>(d.data()), (std::streamsize)len))
        throw error(MAG_ERR_ARG, "short LIME record in " + path);
      return d;
    }
    in.seekg((std::streamoff)padded, std::ios::cur);
  }
}
inline void lime_write(std::ofstream& out, const std::string& type, const void* data, uint64_t len,
                       bool mb, bool me) {
  unsigned char h[144] = {0};
  be_store(h, kLimeMagic, 4);
  be_store(h + 4, 1, 2);
  h[6] = (unsigned char)((mb ? 0x80 : 0) | (me ? 0x40 : 0));
  be_store(h + 8, len, 8);
  std::memcpy(h + 16, type.data(), type.size() < 127 ? type.size() : 127);
  out.write(reinterpret_cast<const char*>(h), sizeof(h));
  out.write(reinterpret_cast<const char*>(data), (std::streamsize)len);
  const char pad[8] = {0};
  out.write(pad, (std::streamsize)((8 - len % 8) % 8));
}
}  // namespace detail

// Readers of the reference's on-disk link formats (data.h:114-176,
// indexing.h:190-351); precision "double" or "float" (ildg: always double).
//   lexicographical: element e of link (site, mu) at (4*site + mu)*N + e
//   qcdstag:         at (site + mu1*V)*N + e', mu1 = (mu + 1) % 4 (t first),
//                    and for su2 the element order reversed: e' = 4 - e, e > 0
//   ildg:            the "ildg-binary-data" record of a LIME file, big-endian
//                    doubles in the lexicographical order (data.h:144-176,
//                    FilePatternILDG indexing.h:325-351); bytes_skip unused
template <class Lattice>
void read_links(Lattice& conf, const std::string& path, const std::string& format,
                long bytes_skip, const std::string& precision) {
  using M = typename Lattice::matrix_type;
  constexpr int N = sizeof(M) / sizeof(double);
  const size_t V = (size_t)conf.lat_dim[0] * conf.lat_dim[1] * conf.lat_dim[2] * conf.lat_dim[3];
  const size_t n = V * 4 * N;
  std::ifstream in(path, std::ios::binary);
  if (!in) throw error(MAG_ERR_ARG, "cannot open " + path);
  std::vector<double> raw(n);
  if (format == "ildg") {
    const std::vector<unsigned char> d = detail::lime_record(in, path, "ildg-binary-data");
    if (d.size() != n * sizeof(double))
      throw error(MAG_ERR_ARG, "ildg-binary-data holds " + std::to_string(d.size()) +
                                   " bytes, the lattice needs " + std::to_string(n * sizeof(double)));
    for (size_t i = 0; i < n; i++) raw[i] = detail::be_double(d.data() + 8 * i);
  } else if (precision == "double") {
    in.ignore(bytes_skip);
    if (!in.read(reinterpret_cast<char*>(raw.data()), n * sizeof(double)))
      throw error(MAG_ERR_ARG, "short read: " + path);
  } else if (precision == "float") {
    in.ignore(bytes_skip);
    std::vector<float> f(n);
    if (!in.read(reinterpret_cast<char*>(f.data()), n * sizeof(float)))
      throw error(MAG_ERR_ARG, "short read: " + path);
    for (size_t i = 0; i < n; i++) raw[i] = f[i];
  } else {
    throw error(MAG_ERR_ARG, "precision must be float or double");
  }
  std::vector<double> rm(n);  // row-major links in the ABI order
  for (size_t s = 0; s < V; s++)
    for (int mu = 0; mu < 4; mu++)
      for (int e = 0; e < N; e++) {
        size_t idx;
        if (format == "lexicographical" || format == "ildg") {
          idx = (s * 4 + mu) * N + e;
        } else if (format == "qcdstag") {
          const int mu1 = (mu == 3) ? 0 : mu + 1;
          const int e1 = (N == 4 && e > 0) ? 4 - e : e;
          idx = (s + (size_t)mu1 * V) * N + e1;
        } else {
          throw error(MAG_ERR_ARG, "unknown format " + format);
        }
        rm[(s * 4 + mu) * N + e] = raw[idx];
      }
  if constexpr (N == 4) {
    std::memcpy(conf.array.data(), rm.data(), n * sizeof(double));
  } else {
    detail::su3_store(conf, rm);
  }
}

// write_data with FilePatternLexicographical (data.h:193-205): doubles
template <class Lattice>
void write_links_lexicographical(const Lattice& conf, const std::string& path) {
  using M = typename Lattice::matrix_type;
  constexpr int N = sizeof(M) / sizeof(double);
  std::ofstream out(path, std::ios::binary);
  if (!out) throw error(MAG_ERR_ARG, "cannot open " + path);
  if constexpr (N == 4) {
    out.write(reinterpret_cast<const char*>(conf.array.data()), conf.array.size() * 4 * sizeof(double));
  } else {
    const std::vector<double> rm = detail::su3_links(conf);
    out.write(reinterpret_cast<const char*>(rm.data()), rm.size() * sizeof(double));
  }
  if (!out) throw error(MAG_ERR_ARG, "write failed: " + path);
}

// ILDG file (a LIME message): an "ildg-format" XML record and the
// "ildg-binary-data" record of big-endian doubles in the lexicographical
// order that read_links(..., "ildg", ...) and the reference read
// (data.h:144-176)
template <class Lattice>
void write_links_ildg(const Lattice& conf, const std::string& path) {
  using M = typename Lattice::matrix_type;
  constexpr int N = sizeof(M) / sizeof(double);
  std::vector<double> rm;
  if constexpr (N == 4) {
    const double* p = reinterpret_cast<const double*>(conf.array.data());
    rm.assign(p, p + conf.array.size() * 4);
  } else {
    rm = detail::su3_links(conf);
  }
  std::vector<unsigned char> be(rm.size() * 8);
  for (size_t i = 0; i < rm.size(); i++) {
    uint64_t u;
    std::memcpy(&u, &rm[i], 8);
    detail::be_store(be.data() + 8 * i, u, 8);
  }
  const std::string xml =
      std::string("<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat>\n<version>1.0</version>\n") +
      "<field>" + (N == 4 ? "su2gauge" : "su3gauge") + "</field>\n<precision>64</precision>\n" +
      "<lx>" + std::to_string(conf.lat_dim[0]) + "</lx><ly>" + std::to_string(conf.lat_dim[1]) +
      "</ly><lz>" + std::to_string(conf.lat_dim[2]) + "</lz><lt>" + std::to_string(conf.lat_dim[3]) +
      "</lt>\n</ildgFormat>";
  std::ofstream out(path, std::ios::binary);
  if (!out) throw error(MAG_ERR_ARG, "cannot open " + path);
  detail::lime_write(out, "ildg-format", xml.data(), xml.size(), true, false);
  detail::lime_write(out, "ildg-binary-data", be.data(), be.size(), false, true);
  if (!out) throw error(MAG_ERR_ARG, "write failed: " + path);
}

}  // namespace mag_amd
