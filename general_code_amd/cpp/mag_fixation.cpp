// mag_fixation -- the MAG fixing app on the MI355X, with the command-line
// flags of ilyakudrov/general_code apps/mag_su2/mag_fixation.cpp:60-117.
//
// For each gauge copy (--gauge_copies, mag_fixation.cpp:165-249): the
// starting spins are new (--is_new_trial 1: uniform random spins, then
// simulated annealing) or read from --path_previous (--is_new_trial 0,
// read_spins); then
//   --is_final 1: final maximisation -> gauge transformation -> MAG
//       functional and plaquette -> lexicographic output of the links, and
//       the copy loop ends;
//   --is_final 0: approximate maximisation; --is_functional_save 1 records
//       the copy's spin functional (written to --path_functional_output as
//       "copy,functional" after the loop); the spins go to
//       --path_spins_output (write_spins), with --is_compare 1 only when
//       their functional beats the old one: that of the spins in
//       --path_previous (--is_compare_spins 1) or of the links in
//       --path_conf read as lexicographical (--is_compare_spins 0).
// The reference's spin-file calls there size the data by process globals it
// never sets and so read and write nothing; here they read and write the
// lattice's V spins (3 doubles each).  Extra flags of this build:
//   --seed N            Philox seed of the spins and the heat bath
//   --synthetic_eps E   no --path_conf: generate exp(i E a.sigma) links with
//                       a Haar gauge transformation from --seed
//   --link_form 1       run the link-form checkerboard omega-overrelaxation
//                       (--omega, --tol, --max_sweeps) instead of the spin
//                       drivers
//   --max_iter N        iteration cap of the maximisation loops
//   --path_monopoles_output P
//                       monopole currents of the fixed links' abelian
//                       projection (calculate_current, monopoles.cpp:669-687)
//                       as 4V doubles J[4*site + mu]
//   --path_output_clusters_unwrapped, --path_output_clusters_wrapped,
//   --path_output_windings, --path_output_monopoles P
//                       the monopole cluster tables of the fixed links, in
//                       the CSV formats of apps/monopoles_su2/
//                       monopoles_su2.cpp:97-197 (clusters on the GPU)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include "mag.h"

using namespace mag_amd;

int main(int argc, char** argv) {
  std::string path_conf, conf_format = "lexicographical", file_precision = "double";
  std::string path_conf_output, path_spins_output, path_previous, path_functional_output;
  std::string path_monopoles_output;
  std::string path_clusters_unwrapped, path_clusters_wrapped, path_windings, path_monopoles;
  int bytes_skip = 0;
  double T_step = 0.01, T_init = 2.5, T_final = 0.01;
  int OR_steps = 6, thermalization_steps = 50;
  double tolerance_maximal = 1e-10, tolerance_average = 1e-14;
  int tolerance_digits = 7, gauge_copies = 1;
  int x_size = 0, y_size = 0, z_size = 0, t_size = 0;
  bool is_new_trial = true, is_final = true, is_functional_save = false;
  bool is_compare = false, is_compare_spins = false;
  uint64_t seed = 1;
  double synthetic_eps = -1;
  bool link_form = false;
  double omega = 1.7, tol = 1e-12;
  int max_sweeps = 1000, max_iter = 1000000;

  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw error(MAG_ERR_ARG, "missing value for " + a);
      return argv[++i];
    };
    if (a == "--conf_format") conf_format = next();
    else if (a == "--file_precision") file_precision = next();
    else if (a == "--path_conf") path_conf = next();
    else if (a == "--path_previous") path_previous = next();
    else if (a == "--path_conf_output") path_conf_output = next();
    else if (a == "--path_spins_output") path_spins_output = next();
    else if (a == "--path_functional_output") path_functional_output = next();
    else if (a == "--bytes_skip") bytes_skip = std::stoi(next());
    else if (a == "--T_step") T_step = std::stod(next());
    else if (a == "--T_init") T_init = std::stod(next());
    else if (a == "--T_final") T_final = std::stod(next());
    else if (a == "--OR_steps") OR_steps = std::stoi(next());
    else if (a == "--tolerance_maximal") tolerance_maximal = std::stod(next());
    else if (a == "--tolerance_average") tolerance_average = std::stod(next());
    else if (a == "--thermalization_steps") thermalization_steps = std::stoi(next());
    else if (a == "--tolerance_digits") tolerance_digits = std::stoi(next());
    else if (a == "--gauge_copies") gauge_copies = std::stoi(next());
    else if (a == "--x_size") x_size = std::stoi(next());
    else if (a == "--y_size") y_size = std::stoi(next());
    else if (a == "--z_size") z_size = std::stoi(next());
    else if (a == "--t_size") t_size = std::stoi(next());
    else if (a == "--is_new_trial") std::istringstream(next()) >> is_new_trial;
    else if (a == "--is_final") std::istringstream(next()) >> is_final;
    else if (a == "--is_functional_save") std::istringstream(next()) >> is_functional_save;
    else if (a == "--is_compare") std::istringstream(next()) >> is_compare;
    else if (a == "--is_compare_spins") std::istringstream(next()) >> is_compare_spins;
    else if (a == "--ml5_conf_num") next();  // parsed and unused, as in the reference
    else if (a == "--seed") seed = std::stoull(next());
    else if (a == "--synthetic_eps") synthetic_eps = std::stod(next());
    else if (a == "--link_form") std::istringstream(next()) >> link_form;
    else if (a == "--path_monopoles_output") path_monopoles_output = next();
    else if (a == "--path_output_clusters_unwrapped") path_clusters_unwrapped = next();
    else if (a == "--path_output_clusters_wrapped") path_clusters_wrapped = next();
    else if (a == "--path_output_windings") path_windings = next();
    else if (a == "--path_output_monopoles") path_monopoles = next();
    else if (a == "--omega") omega = std::stod(next());
    else if (a == "--tol") tol = std::stod(next());
    else if (a == "--max_sweeps") max_sweeps = std::stoi(next());
    else if (a == "--max_iter") max_iter = std::stoi(next());
    else {
      std::cerr << "unknown parameter " << a << std::endl;
      return 2;
    }
  }
  try {
    const std::array<int, 4> dims = {x_size, y_size, z_size, t_size};
    Data::LatticeData<DataPatternLexicographical, su2> conf(dims);
    if (!path_conf.empty()) {
      read_links(conf, path_conf, conf_format, bytes_skip, file_precision);
    } else if (synthetic_eps >= 0) {
      mag_params p;
      mag_params_default(&p);
      for (int k = 0; k < 4; k++) p.dims[k] = dims[k];
      mag_job* job = nullptr;
      check(mag_job_create(&p, &job));
      int rc = mag_job_generate(job, synthetic_eps, seed);
      if (rc == MAG_OK) rc = mag_job_download_links(job, detail::su2_links(conf));
      mag_job_destroy(job);
      check(rc);
    } else {
      std::cerr << "need --path_conf or --synthetic_eps" << std::endl;
      return 2;
    }
    std::cout.precision(17);
    std::cout << "plaquette " << plaket(conf) << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<double> J;
    const bool want_clusters = !path_clusters_unwrapped.empty() || !path_clusters_wrapped.empty() ||
                               !path_windings.empty() || !path_monopoles.empty();
    const bool want_J = !path_monopoles_output.empty() || want_clusters;
    bool fixed = false;  // the links hold the gauge-fixed configuration
    if (link_form) {
      fixed = true;
      std::vector<double> hist;
      const int sweeps = fix_mag_links(conf, omega, tol, max_sweeps, MAG_PREC_F64, &hist,
                                       want_J ? &J : nullptr);
      std::cout << "link-form sweeps " << sweeps << std::endl;
    } else {
      DataPatternLexicographical pattern(dims);
      std::vector<spin> spins;
      std::map<int, double> functional;
      if (!is_final && path_spins_output.empty())
        throw error(MAG_ERR_ARG, "--is_final 0 writes the spins: give --path_spins_output");
      if (!is_new_trial && path_previous.empty())
        throw error(MAG_ERR_ARG, "--is_new_trial 0 reads the spins from --path_previous");
      for (int copy = 1; copy <= gauge_copies; copy++) {
        if (is_new_trial) {
          spins = generate_spins_uniform<spin>(pattern, seed + copy);
          make_simulated_annealing(conf, spins, T_init, T_final, T_step, OR_steps,
                                   thermalization_steps, seed + copy);
        } else {
          spins = read_spins<spin>(path_previous, pattern);
        }
        if (is_final) {
          const int it = make_maximization_final(conf, spins, OR_steps, tolerance_maximal,
                                                 tolerance_average, max_iter);
          std::cout << "maximization final iterations " << it << std::endl;
          std::cout << "spin functional " << MAG_functional_su2_spin(conf, spins) << std::endl;
          gauge_tranformation_spins(conf, spins);
          fixed = true;
          break;
        }
        const int it = make_maximization_approximate(conf, spins, OR_steps, tolerance_digits,
                                                     max_iter);
        const double F_new = MAG_functional_su2_spin(conf, spins);
        std::cout << "copy " << copy << " maximization approximate iterations " << it
                  << std::endl;
        std::cout << "spin functional " << F_new << std::endl;
        if (is_functional_save) functional[copy] = F_new;
        if (is_compare) {
          double F_old;
          if (is_compare_spins) {
            std::vector<spin> old = read_spins<spin>(path_previous, pattern);
            F_old = MAG_functional_su2_spin(conf, old);
          } else {
            Data::LatticeData<DataPatternLexicographical, su2> conf_old(dims);
            read_links(conf_old, path_conf, "lexicographical", 0, file_precision);
            F_old = MAG_functional_su2(conf_old);
          }
          std::cout << "new functional is " << F_new << std::endl;
          std::cout << "old functional is " << F_old << std::endl;
          if (F_old < F_new) {
            std::cout << "new functional is higher, saving spin configuration" << std::endl;
            write_spins(path_spins_output, spins);
          } else {
            std::cout << "new functional is lower, spin configuration is not saved" << std::endl;
          }
        } else {
          write_spins(path_spins_output, spins);
        }
      }
      if (is_functional_save) {
        std::ofstream f(path_functional_output);
        if (!f) throw error(MAG_ERR_ARG, "cannot write " + path_functional_output);
        f.precision(17);
        f << "copy,functional" << std::endl;
        for (const auto& [c, F] : functional) f << c << "," << F << std::endl;
      }
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "fixing time " << secs << std::endl;
    if (!fixed) return 0;  // approximate path: the spins are the output
    std::cout << "final functional is " << MAG_functional_su2(conf) << std::endl;
    std::cout << "plaquette " << plaket(conf) << std::endl;
    if (!path_conf_output.empty()) write_links_lexicographical(conf, path_conf_output);
    if (want_J) {
      if (J.empty()) {
        // spin path: the fixed links are on the host; project them as the
        // monopole app reads them (MatrixConverter<su2, abelian>)
        Data::LatticeData<DataPatternLexicographical, abelian> ab(dims);
        for (size_t l = 0; l < conf.array.size(); l++)
          ab.array[l] = abelian{1.0, std::atan2(conf.array[l].a3, conf.array[l].a0)};
        J = calculate_current(ab);
      }
      if (!path_monopoles_output.empty()) {
        std::ofstream f(path_monopoles_output, std::ios::binary);
        f.write(reinterpret_cast<const char*>(J.data()), (std::streamsize)(J.size() * sizeof(double)));
        if (!f) throw error(MAG_ERR_ARG, "cannot write " + path_monopoles_output);
      }
      long n = 0;
      for (double v : J) n += (long)std::lround(std::fabs(v));
      std::cout << "monopole current links " << n << std::endl;
    }
    if (want_clusters) {
      DataPatternLexicographical pattern(dims);
      std::vector<loop_new*> LL = calculate_clusters(J, pattern);
      const cluster_tables t = make_cluster_tables(LL, dims);
      std::cout << "monopole clusters " << LL.size() << std::endl;
      free_clusters(LL);
      auto open = [](const std::string& path) {
        std::ofstream f(path);
        if (!f) throw error(MAG_ERR_ARG, "cannot write " + path);
        return f;  // default stream precision, as the reference app's files
      };
      if (!path_clusters_unwrapped.empty()) {
        auto f = open(path_clusters_unwrapped);
        f << "length,number" << std::endl;
        for (const auto& [len, num] : t.unwrapped) f << len << "," << num << std::endl;
      }
      if (!path_clusters_wrapped.empty()) {
        auto f = open(path_clusters_wrapped);
        f << "length,x0_wrap,x1_wrap,x2_wrap,x3_wrap,percolating_group" << std::endl;
        for (size_t i = 0; i < t.wrapped_lengths.size(); i++) {
          const bool perc = std::find(t.percolating.begin(), t.percolating.end(), (int)i) != t.percolating.end();
          f << t.wrapped_lengths[i];
          for (int w : t.wrappings[i]) f << "," << w;
          f << (perc ? ",percolating" : ",non-percolating") << std::endl;
        }
      }
      if (!path_windings.empty()) {
        // the reference app writes the header only: its winding maps are
        // never filled (monopoles_su2.cpp:122-123, 186-191)
        auto f = open(path_windings);
        f << "winding_number,cluster_number,direction" << std::endl;
      }
      if (!path_monopoles.empty()) {
        auto f = open(path_monopoles);
        f << "asymmetry" << std::endl << t.asymmetry << std::endl;
      }
    }
  } catch (const error& e) {
    std::cerr << "mag_fixation: " << e.what() << " (status " << e.code << ")" << std::endl;
    return 1;
  } catch (const std::exception& e) {
    std::cerr << "mag_fixation: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
