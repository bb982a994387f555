// illum_cli: command-line driver over the GPU library, with the reference
// driver's tasks and flags (illum_cli.cpp:270-510 tasks, :561-711 flags,
// :34-160 YAML keys).
//
//   illum_cli form_factors -p mesh.obj [--output_dir d]          -> d/form_factors.bin
//   illum_cli horizons     -p mesh.obj [--nphi 361 --theta_eps 1e-3 --offset 1e-5]
//                                                                 -> d/horizons.bin
//   illum_cli radiosity    -p mesh.obj [-r] [-t] [--sun_pos_file f --sun_unit m]
//                          [--albedo 0.12|path] [--dt 600 --ti 70 --rhoc .. --depths ..
//                           --T0 233|path] [--repeat n] [--gs_steps n --print_residual]
//                                                                 -> rad.bin, rad_avg.bin,
//                                                                    therm*.bin
// Extra (GPU) flags: --synth s[,craters[,seed]] (seeded cratered ellipsoid
// instead of an OBJ), --precision fp64|mixed, --rad_tol, --leaf_size.
//
// Reference behaviours kept on purpose (documented quirks):
//   * compute_F is called with the default offset 1e-5 whatever --offset says
//     (illum_cli.cpp:275, 392);
//   * the frame counter is never incremented, so rad_avg / therm_avg hold the
//     last frame and therm_max stays 0 (illum_cli.cpp:401, 442-453);
//   * without --sun_pos_file one sun position is used at 227,390,024 km
//     (illum_cli.cpp:336-340; here in a fixed direction instead of randn).
// The scattered-radiosity solve runs to --rad_tol (default 1e-10 relative)
// instead of a fixed number of Gauss-Seidel sweeps; --gs_steps caps the number
// of sparse products at 4 x gs_steps + 8 when given.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/illum_hip.hpp"
#include "arma_io.hpp"

namespace {

using Args = std::map<std::string, std::string>;

const char *USAGE =
    "usage: illum_cli <form_factors|horizons|radiosity> -p mesh.obj [options]\n"
    "  --config f.yaml  --offset 1e-5  --theta_eps 1e-3  --nphi 361  -r/--radiosity\n"
    "  --gs_steps n  --print_residual  -t/--thermal  -q/--quiet  --dt 600  --albedo 0.12|path\n"
    "  --depths a,b,..  --ti 70  --rhoc a[,b,..]  --T0 233|path  --output_dir d\n"
    "  --form_factors_file f  --horizon_file f  --horizon_obj_file f  --sun_pos_file f\n"
    "  --sun_unit m|km  --save_full_thermal  --record_therm_stats  --repeat n  --print_config\n"
    "  --synth s[,craters[,seed]]  --precision fp64|mixed  --rad_tol 1e-10  --leaf_size 4\n";

bool is_bool_flag(const std::string &k) {
  static const char *b[] = {"radiosity", "thermal", "quiet", "print_residual", "save_full_thermal",
                            "record_therm_stats", "print_config", "help"};
  for (auto *x : b)
    if (k == x) return true;
  return false;
}

// Flat YAML subset used by share/config.yaml: "key: value" and "key: [a, b, ...]"
// (lists may span lines); '#' comments.
void read_yaml(const std::string &path, Args &a) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot read config " + path);
  std::string line, key, acc;
  bool in_list = false;
  while (std::getline(f, line)) {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    if (in_list) {
      acc += line;
      if (line.find(']') != std::string::npos) {
        in_list = false;
      } else {
        continue;
      }
    } else {
      const auto colon = line.find(':');
      if (colon == std::string::npos) continue;
      key = line.substr(0, colon);
      key.erase(0, key.find_first_not_of(" \t"));
      key.erase(key.find_last_not_of(" \t") + 1);
      acc = line.substr(colon + 1);
      if (acc.find('[') != std::string::npos && acc.find(']') == std::string::npos) {
        in_list = true;
        continue;
      }
    }
    std::string v;
    for (char c : acc)
      if (c != '[' && c != ']' && c != ' ' && c != '\t') v += c;
    if (key == "thermal_inertia") key = "ti";
    if (!v.empty() && !a.count(key)) a[key] = v;  // command-line flags win
  }
}

std::vector<double> parse_list(const std::string &s) {
  std::vector<double> v;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ','))
    if (!tok.empty()) v.push_back(std::stod(tok));
  return v;
}

bool flag(const Args &a, const std::string &k) {
  auto it = a.find(k);
  return it != a.end() && it->second != "false" && it->second != "0";
}

std::string get(const Args &a, const std::string &k, const std::string &def) {
  auto it = a.find(k);
  return it == a.end() ? def : it->second;
}

template <class F>
void timed(const std::string &msg, F &&fn) {
  std::cout << msg << std::flush;
  const auto t0 = std::chrono::steady_clock::now();
  fn();
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << " [" << s << "s]" << std::endl;
}

bool is_number(const std::string &s) {
  char *end = nullptr;
  std::strtod(s.c_str(), &end);
  return !s.empty() && end && *end == '\0';
}

}  // namespace

int main(int argc, char **argv) {
  try {
    Args a;
    std::string task = "radiosity";
    bool task_set = false;
    for (int i = 1; i < argc; ++i) {
      std::string s = argv[i];
      if (s == "-h" || s == "--help") { std::cout << USAGE; return 0; }
      if (s == "-r") { a["radiosity"] = "true"; continue; }
      if (s == "-t") { a["thermal"] = "true"; continue; }
      if (s == "-q") { a["quiet"] = "true"; continue; }
      if (s == "-p") { if (i + 1 >= argc) throw std::runtime_error("-p needs a path"); a["path"] = argv[++i]; continue; }
      if (s.rfind("--", 0) == 0) {
        std::string k = s.substr(2), v;
        const auto eq = k.find('=');
        if (eq != std::string::npos) { v = k.substr(eq + 1); k = k.substr(0, eq); }
        else if (is_bool_flag(k)) v = "true";
        else { if (i + 1 >= argc) throw std::runtime_error("--" + k + " needs a value"); v = argv[++i]; }
        a[k] = v;
        continue;
      }
      if (!task_set) { task = s; task_set = true; continue; }
      throw std::runtime_error("unexpected argument " + s);
    }
    if (a.count("task")) task = a["task"];
    if (a.count("config")) read_yaml(a["config"], a);
    if (flag(a, "print_config"))
      for (auto &kv : a) std::cout << kv.first << ": " << kv.second << "\n";

    const std::string out_dir = get(a, "output_dir", ".");
    const double offset = std::stod(get(a, "offset", "1e-5"));
    const double theta_eps = std::stod(get(a, "theta_eps", "1e-3"));
    const int nphi = std::stoi(get(a, "nphi", "361"));
    const int leaf = std::stoi(get(a, "leaf_size", "4"));
    const bool quiet = flag(a, "quiet");

    illum::context gpu(0);
    std::unique_ptr<illum::illum_context> ctx;
    timed("- loading mesh", [&] {
      if (a.count("synth")) {
        std::vector<double> s = parse_list(a["synth"]);
        const int subdiv = (int)s.at(0), ncr = s.size() > 1 ? (int)s[1] : 40;
        const uint64_t seed = s.size() > 2 ? (uint64_t)s[2] : 0;
        int64_t nv, nf;
        illum::check(illum_synth_sizes(subdiv, &nv, &nf));
        std::vector<double> V(3 * nv);
        std::vector<int32_t> F(3 * nf);
        illum::check(illum_synth_ellipsoid(subdiv, 13, 11, 9, ncr, 0.05, 0.3, 0.2, 1.5, seed,
                                           V.data(), F.data()));
        ctx.reset(new illum::illum_context(gpu, V, F, leaf));
      } else {
        if (!a.count("path")) throw std::runtime_error("no mesh: give -p mesh.obj or --synth s");
        ctx.reset(new illum::illum_context(gpu, a["path"], leaf));
      }
      std::cout << " [" << ctx->get_num_faces() << " faces]";
    });
    if (a.count("horizon_obj_file")) ctx->set_horizon_obj_file(a["horizon_obj_file"]);
    const int64_t nf = ctx->num_faces;

    // albedo: a number, or a path to an Armadillo vector (illum.cpp:225-234)
    auto albedo_vec = [&]() -> std::vector<double> {
      const std::string s = get(a, "albedo", "0.12");
      if (is_number(s)) return {};
      int64_t r, c;
      return illum_cli::load_dense(s, &r, &c);
    };
    auto form_factors = [&]() {
      const std::string s = get(a, "albedo", "0.12");
      auto av = albedo_vec();
      // compute_F is called without --offset in the reference (illum_cli.cpp:275, 392)
      return av.empty() ? ctx->compute_F(std::stod(s)) : ctx->compute_F(av);
    };

    if (task == "form_factors") {
      illum::sp_mat F;
      timed("- assembling form factor matrix", [&] { F = form_factors(); std::cout << " [nnz = " << F.n_nonzero() << "]"; });
      timed("- saving form factor matrix", [&] {
        std::vector<int64_t> cp; std::vector<int32_t> ri; std::vector<double> v;
        F.csc(cp, ri, v);
        illum_cli::save_sparse(out_dir + "/form_factors.bin", nf, nf, cp, ri, v);
      });
      return 0;
    }
    if (task == "horizons") {
      timed("- building horizon map", [&] { ctx->make_horizons(nphi, theta_eps, offset); });
      if (!quiet)
        timed("- saving horizon map to " + out_dir + "/horizons", [&] {
          ctx->save_horizons(out_dir + "/horizons.bin");
        });
      return 0;
    }
    if (task != "radiosity") throw std::runtime_error("unknown task " + task);

    // ---- radiosity task (illum_cli.cpp:305-510) ----
    int hz_nphi = nphi;
    if (a.count("horizon_file")) {
      timed("- loading horizon map from " + a["horizon_file"], [&] {
        ctx->load_horizons(a["horizon_file"]);
        hz_nphi = ctx->nphi();
      });
    } else {
      timed("- building horizon map", [&] { ctx->make_horizons(nphi, theta_eps, offset); });
    }
    std::vector<double> sun;  // 3 x nsun column-major
    if (a.count("sun_pos_file")) {
      timed("- loading sun positions", [&] {
        for (auto &row : illum_cli::load_raw_ascii(a["sun_pos_file"])) {
          if (row.size() < 3) throw std::runtime_error("sun position rows need 3 columns");
          sun.insert(sun.end(), row.begin(), row.begin() + 3);
        }
      });
    } else {
      const double d = 227390024000.;  // m
      const double u[3] = {0.6, 0.48, 0.64};
      for (double x : u) sun.push_back(d * x);
    }
    if (get(a, "sun_unit", "m") == "m")
      for (auto &x : sun) x /= 1000.;
    const int nsun = (int)(sun.size() / 3);
    const std::vector<double> disk_xy = illum::fib_spiral(100);

    const bool rad_on = flag(a, "radiosity"), therm_on = flag(a, "thermal");
    const int prec_therm = get(a, "precision", "fp64") == "mixed" ? ILLUM_THERM_MIXED : ILLUM_THERM_FP64;
    const int method = get(a, "precision", "fp64") == "mixed" ? ILLUM_RAD_MIXED : ILLUM_RAD_JACOBI_FP64;
    const double rad_tol = std::stod(get(a, "rad_tol", "1e-10"));
    const int max_products = a.count("gs_steps") ? 4 * std::stoi(a["gs_steps"]) + 8 : 400;

    std::unique_ptr<illum::thermal_model> therm;
    if (therm_on) {
      std::vector<double> depths = a.count("depths") ? parse_list(a["depths"]) : std::vector<double>{};
      if (depths.size() < 2) throw std::runtime_error("--thermal needs --depths with >= 2 layers");
      std::vector<double> rhoc = a.count("rhoc") ? parse_list(a["rhoc"]) : std::vector<double>{1.19177e6};
      const std::string T0 = get(a, "T0", "233.0");
      if (!is_number(T0)) throw std::runtime_error("--T0 from a file: use the library API (illum_thermal_create with T0_full)");
      therm.reset(new illum::thermal_model(gpu, (int)nf, depths, std::stod(get(a, "ti", "70.0")),
                                           rhoc, std::stod(T0), 0.999, prec_therm));
    }
    illum::sp_mat F;
    if (rad_on) {
      if (a.count("form_factors_file")) {
        timed("- loading form factor matrix", [&] {
          int64_t r, c; std::vector<int64_t> cp; std::vector<int32_t> ri; std::vector<double> v;
          illum_cli::load_sparse(a["form_factors_file"], &r, &c, cp, ri, v);
          // CSC -> CSR on the host (file I/O path only)
          std::vector<int64_t> rp(r + 1, 0);
          for (auto i : ri) rp[i + 1]++;
          for (int64_t i = 0; i < r; ++i) rp[i + 1] += rp[i];
          std::vector<int32_t> ci(v.size()); std::vector<double> cv(v.size());
          std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
          for (int64_t j = 0; j < c; ++j)
            for (int64_t k = cp[j]; k < cp[j + 1]; ++k) { ci[cur[ri[k]]] = (int32_t)j; cv[cur[ri[k]]++] = v[k]; }
          F = illum::sp_mat::from_csr(gpu, r, rp, ci, cv);
          std::cout << " [nnz = " << F.n_nonzero() << "]";
        });
      } else {
        timed("- assembling form factor matrix", [&] { F = form_factors(); std::cout << " [nnz = " << F.n_nonzero() << "]"; });
      }
    }

    std::vector<double> rad(nf, 0.0), rad_avg(nf, 0.0), therm_v, therm_avg, therm_max;
    if (therm_on) { therm_v.assign(nf, 0.0); therm_avg.assign(nf, 0.0); therm_max.assign(nf, 0.0); }
    std::vector<std::vector<double>> stats;
    const int current_frame = 0;  // never incremented in the reference (illum_cli.cpp:401)
    const double dt = std::stod(get(a, "dt", "600"));
    const bool print_res = flag(a, "print_residual");

    auto step = [&](int j, int rep) {
      std::string frame = std::to_string(j + 1) + "/" + std::to_string(nsun);
      if (rep > 0) frame += " [" + std::to_string(rep) + "/" + get(a, "repeat", "0") + "]";
      std::fill(rad.begin(), rad.end(), 0.0);
      if (therm_on) {
        auto Fs = therm->get_radiosity();
        for (int64_t i = 0; i < nf; ++i) rad[i] += Fs[i];
      }
      timed("- " + frame + ": computing direct radiosity", [&] {
        std::vector<double> s3(sun.begin() + 3 * j, sun.begin() + 3 * j + 3);
        // direct radiosity against the loaded/computed horizon map
        const std::vector<double> E = ctx->get_direct_radiosity(s3, disk_xy);
        double mean = 0;
        for (int64_t i = 0; i < nf; ++i) { rad[i] += E[i]; mean += rad[i]; }
        std::cout << " [mean: " << mean / nf << " W/m^2]";
      });
      if (rad_on) {
        timed("  + solving for scattered radiosity", [&] {
          auto r = illum::solve_radiosity(F, rad, 1, method, rad_tol, max_products);
          rad = r.B;
          if (print_res) {
            int n = 0;
            for (double x : r.residuals) if (x != 0) ++n;
            std::cout << " [rel res = " << (n ? r.residuals[n - 1] : 0.0) << ", products = " << r.products << "]";
          }
        });
      }
      for (int64_t i = 0; i < nf; ++i) rad_avg[i] += (rad[i] - rad_avg[i]) / (current_frame + 1);
      if (therm_on) {
        timed("- " + frame + ": stepping thermal model", [&] {
          therm->step(dt, rad);
          auto T = therm->T();
          const int nzp = therm->nz + 1;
          for (int64_t i = 0; i < nf; ++i) therm_v[i] = T[(size_t)i * nzp];
          for (int64_t i = 0; i < nf; ++i) therm_avg[i] += (therm_v[i] - therm_avg[i]) / (current_frame + 1);
          if (current_frame > 50)
            for (int64_t i = 0; i < nf; ++i) therm_max[i] = std::max(therm_max[i], therm_v[i]);
          if (flag(a, "record_therm_stats")) {
            // Tmax, Tmin, Tavg, Tstd over all / top / bottom (thermal.hpp:86-104)
            std::vector<double> row;
            for (int part = 0; part < 3; ++part) {
              double mx = -1e300, mn = 1e300, s = 0, s2 = 0;
              int64_t cnt = 0;
              for (int64_t i = 0; i < nf; ++i)
                for (int l = 0; l < nzp; ++l) {
                  if ((part == 1 && l != 0) || (part == 2 && l != nzp - 1)) continue;
                  const double x = T[(size_t)i * nzp + l];
                  mx = std::max(mx, x); mn = std::min(mn, x); s += x; ++cnt;
                }
              const double mean = s / cnt;
              for (int64_t i = 0; i < nf; ++i)
                for (int l = 0; l < nzp; ++l) {
                  if ((part == 1 && l != 0) || (part == 2 && l != nzp - 1)) continue;
                  const double x = T[(size_t)i * nzp + l] - mean;
                  s2 += x * x;
                }
              row.insert(row.end(), {mx, mn, mean, std::sqrt(s2 / (cnt > 1 ? cnt - 1 : 1))});
            }
            stats.push_back(row);
          }
        });
      }
    };
    if (a.count("repeat")) {
      const int rep = std::stoi(a["repeat"]);
      for (int r = 1; r <= rep; ++r)
        for (int j = 0; j < nsun; ++j) step(j, r);
    } else {
      for (int j = 0; j < nsun; ++j) step(j, 0);
    }
    (void)hz_nphi;
    if (!quiet) {
      timed("- saving radiosity", [&] { illum_cli::save_dense(out_dir + "/rad.bin", rad, nf, 1); });
      timed("- saving average radiosity", [&] { illum_cli::save_dense(out_dir + "/rad_avg.bin", rad_avg, nf, 1); });
      if (therm_on) {
        if (flag(a, "save_full_thermal")) {
          timed("- saving full thermal model", [&] {
            illum_cli::save_dense(out_dir + "/therm.bin", therm->T(), therm->nz + 1, nf);
          });
        } else {
          timed("- saving thermal", [&] { illum_cli::save_dense(out_dir + "/therm.bin", therm_v, nf, 1); });
        }
        timed("- saving average thermal", [&] { illum_cli::save_dense(out_dir + "/therm_avg.bin", therm_avg, nf, 1); });
        timed("- saving max thermal", [&] { illum_cli::save_dense(out_dir + "/therm_max.bin", therm_max, nf, 1); });
      }
      if (flag(a, "record_therm_stats") && !stats.empty()) {
        timed("- saving thermal stats", [&] {
          std::vector<double> cm(stats.size() * 12);
          for (size_t k = 0; k < stats.size(); ++k)
            for (int c = 0; c < 12; ++c) cm[c * stats.size() + k] = stats[k][c];
          illum_cli::save_dense(out_dir + "/therm_stats.bin", cm, (int64_t)stats.size(), 12);
        });
      }
    }
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "illum_cli: " << e.what() << std::endl;
    return 1;
  }
}
