// Plugin benchmark: loads libblockdaglu.so at run time (dlopen + dlsym of
// CreateAlgorithm / DestroyAlgorithm, the way the reference's dll_benchmark
// loads its DLL, apps/dll_benchmark/main.cpp:209-215), drives the ALInterface
// stages with stage timers (apps/benchmark/main.cpp:137-198) and writes the
// reference's evidence files with the same columns (:309-432):
//   runs.csv            alg,case,repetition,strategy,n,nnz,reorder_ms,lu_ms,solve_ms,total_ms,rel_err,rel_residual,ok
//   margin.csv          case,n,nnz,xstar_rel_residual,ss_rel_residual,ua_rel_residual,eps_suitesparse,err_user,
//                       margin_eps_minus_user,pass_user_le_eps
//   timing_summary.csv  alg,case,strategy,n,nnz,repetitions,median_total_ms,p95_total_ms,median_rel_err,
//                       median_rel_residual,ok
//   summary.txt         <alg>_weighted_total_ms, <alg>_max_rel_err, <alg>_all_ok,
//                       speedup_suitesparse_over_blockdaglu
// Cases: <root>/<id>/<id>_matrix.coo ("n nnz" then "row col value" lines,
// 0-based), <id>_rhs.csv and <id>_xstar.csv (one value per line). The baseline
// column (named "SuiteSparse" in the reference, whose baseline is the
// organizer's KLU) comes from <root>/baseline_runs.csv in the runs.csv format,
// written by whoever ran the baseline; without it the margin gate is skipped.
// Usage: bdlu_plugin_bench --lib <libblockdaglu.so> <cases_root> [case_id]
//        bdlu_plugin_bench --lib <libblockdaglu.so> --selftest
// Env: BLOCKDAGLU_BENCH_OUT (output directory, else stdout),
//      BLOCKDAGLU_BENCH_WARMUP (1), BLOCKDAGLU_BENCH_REPETITIONS (10).
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "ALInterface.h"

namespace fs = std::filesystem;
using Clock = std::chrono::steady_clock;

namespace {

using CreateFn = ALInterface* (*)();
using DestroyFn = void (*)(ALInterface*);
using ModeFn = const char* (*)(ALInterface*);

struct Plugin {
  CreateFn create = nullptr;
  DestroyFn destroy = nullptr;
  ModeFn mode = nullptr;
};

struct Run {
  std::string alg, strategy;
  int case_id = 0, repetition = 0, n = 0;
  std::size_t nnz = 0;
  double reorder_ms = 0, lu_ms = 0, solve_ms = 0, total_ms = 0, rel_err = 0, rel_residual = 0;
  bool ok = false;
};

constexpr double kInf = std::numeric_limits<double>::infinity();

std::vector<double> read_column(const fs::path& p) {
  std::vector<double> v;
  std::ifstream f(p);
  std::string line;
  while (std::getline(f, line))
    if (!line.empty()) v.push_back(std::stod(line));
  return v;
}

bool read_coo(const fs::path& p, int& n, COOMatrix& a) {
  std::ifstream f(p);
  std::size_t nnz = 0;
  if (!(f >> n >> nnz)) return false;
  a.clear();
  a.reserve(nnz);
  int i, j;
  double v;
  while (a.size() < nnz && (f >> i >> j >> v)) a.emplace_back(i, j, v);
  return a.size() == nnz;
}

double rel_l2(const std::vector<double>& x, const std::vector<double>& ref) {
  if (x.size() != ref.size() || ref.empty()) return kInf;
  double num = 0, den = 0;
  for (std::size_t i = 0; i < x.size(); ++i) {
    num += (x[i] - ref[i]) * (x[i] - ref[i]);
    den += ref[i] * ref[i];
  }
  return den > 0 ? std::sqrt(num / den) : kInf;
}

double rel_residual(const COOMatrix& a, int n, const std::vector<double>& x, const std::vector<double>& b) {
  if (static_cast<int>(x.size()) != n || static_cast<int>(b.size()) != n) return kInf;
  std::vector<double> r(b);
  for (const auto& [i, j, v] : a) r[i] -= v * x[j];
  double num = 0, den = 0;
  for (int i = 0; i < n; ++i) {
    num += r[i] * r[i];
    den += b[i] * b[i];
  }
  return den > 0 ? std::sqrt(num / den) : kInf;
}

Run run_case(const Plugin& pl, int id, int repetition, const COOMatrix& a, int n, const std::vector<double>& b0,
             const std::vector<double>& xstar) {
  Run r;
  r.alg = "BlockDAGLU";
  r.case_id = id;
  r.repetition = repetition;
  r.n = n;
  r.nnz = a.size();
  ALInterface* alg = pl.create();
  if (!alg) return r;
  std::vector<double> b(b0);
  COOMatrix copy(a);
  if (!alg->prepareData(std::move(copy))) {
    pl.destroy(alg);
    return r;
  }
  const auto t0 = Clock::now();
  const bool ok_re = alg->rearrangeMatrix();
  const auto t1 = Clock::now();
  if (ok_re && pl.mode) r.strategy = pl.mode(alg);
  const bool ok_lu = ok_re && alg->decompositionLU();
  const auto t2 = Clock::now();
  const bool ok_sol = ok_lu && alg->solveLinearSystem(b);
  const auto t3 = Clock::now();
  pl.destroy(alg);
  auto ms = [](Clock::time_point x, Clock::time_point y) { return std::chrono::duration<double, std::milli>(y - x).count(); };
  r.reorder_ms = ms(t0, t1);
  r.lu_ms = ms(t1, t2);
  r.solve_ms = ms(t2, t3);
  r.total_ms = r.reorder_ms + r.lu_ms + r.solve_ms;
  r.ok = ok_sol;
  r.rel_err = ok_sol ? rel_l2(b, xstar) : kInf;
  r.rel_residual = ok_sol ? rel_residual(a, n, b, b0) : kInf;
  return r;
}

std::vector<const Run*> runs_of(const std::vector<Run>& all, const std::string& alg, int id) {
  std::vector<const Run*> m;
  for (const Run& r : all)
    if (r.alg == alg && r.case_id == id) m.push_back(&r);
  std::sort(m.begin(), m.end(), [](const Run* x, const Run* y) { return x->total_ms < y->total_ms; });
  return m;
}
const Run* median(const std::vector<Run>& all, const std::string& alg, int id) {
  const auto m = runs_of(all, alg, id);
  return m.empty() ? nullptr : m[m.size() / 2];
}
double p95(const std::vector<Run>& all, const std::string& alg, int id) {
  const auto m = runs_of(all, alg, id);
  if (m.empty()) return kInf;
  const std::size_t k = static_cast<std::size_t>(std::ceil(0.95 * m.size()) - 1.0);
  return m[std::min(k, m.size() - 1)]->total_ms;
}
double weight_for_n(int n) {  // apps/benchmark/main.cpp:200-205
  return n <= 20000 ? 0.1 : n <= 40000 ? 0.15 : n <= 55000 ? 0.25 : 0.5;
}

std::vector<Run> read_baseline(const fs::path& p) {
  std::vector<Run> out;
  std::ifstream f(p);
  std::string line;
  if (!std::getline(f, line)) return out;  // header
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    std::vector<std::string> c;
    std::string cell;
    while (std::getline(ss, cell, ',')) c.push_back(cell);
    if (c.size() < 13) continue;
    Run r;
    r.alg = c[0];
    r.case_id = std::stoi(c[1]);
    r.repetition = std::stoi(c[2]);
    r.strategy = c[3];
    r.n = std::stoi(c[4]);
    r.nnz = std::stoull(c[5]);
    r.reorder_ms = std::stod(c[6]);
    r.lu_ms = std::stod(c[7]);
    r.solve_ms = std::stod(c[8]);
    r.total_ms = std::stod(c[9]);
    r.rel_err = std::stod(c[10]);
    r.rel_residual = std::stod(c[11]);
    r.ok = c[12] == "1";
    out.push_back(r);
  }
  return out;
}

int env_int(const char* name, int def) {
  const char* v = std::getenv(name);
  const int x = v ? std::atoi(v) : 0;
  return x > 0 ? x : def;
}

}  // namespace

int main(int argc, char** argv) {
  std::string lib;
  std::vector<std::string> pos;
  bool selftest = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--lib") && i + 1 < argc) lib = argv[++i];
    else if (!std::strcmp(argv[i], "--selftest")) selftest = true;
    else pos.push_back(argv[i]);
  }
  if (lib.empty() || (!selftest && pos.empty())) {
    std::cerr << "usage: bdlu_plugin_bench --lib <libblockdaglu.so> (<cases_root> [case_id] | --selftest)\n";
    return 2;
  }
  void* h = dlopen(lib.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    std::cerr << "dlopen failed: " << dlerror() << "\n";
    return 6;
  }
  Plugin pl;
  pl.create = reinterpret_cast<CreateFn>(dlsym(h, "CreateAlgorithm"));
  pl.destroy = reinterpret_cast<DestroyFn>(dlsym(h, "DestroyAlgorithm"));
  pl.mode = reinterpret_cast<ModeFn>(dlsym(h, "AlgorithmSelectedMode"));
  if (!pl.create || !pl.destroy) {
    std::cerr << "CreateAlgorithm / DestroyAlgorithm not exported by " << lib << "\n";
    return 7;
  }
  if (selftest) {
    // Host-side stages only: create, store a matrix, reject a bad one, destroy.
    ALInterface* a = pl.create();
    const bool stored = a && a->prepareData({{0, 0, 2.0}, {1, 1, 3.0}});
    ALInterface* bad = pl.create();
    const bool rejected = bad && bad->prepareData({{-1, 0, 1.0}}) && !bad->rearrangeMatrix();
    if (a) pl.destroy(a);
    if (bad) pl.destroy(bad);
    std::cout << "plugin selftest: created=" << (a != nullptr) << " stored=" << stored << " rejected=" << rejected << "\n";
    return stored && rejected ? 0 : 1;
  }

  const fs::path root(pos[0]);
  const int only = pos.size() > 1 ? std::atoi(pos[1].c_str()) : 0;
  std::vector<int> ids;
  for (const auto& e : fs::directory_iterator(root)) {
    const std::string nm = e.path().filename().string();
    if (e.is_directory() && !nm.empty() && std::all_of(nm.begin(), nm.end(), ::isdigit)) ids.push_back(std::stoi(nm));
  }
  std::sort(ids.begin(), ids.end());
  if (only > 0) ids.erase(std::remove_if(ids.begin(), ids.end(), [&](int x) { return x != only; }), ids.end());
  if (ids.empty()) {
    std::cerr << "no numeric case folders under " << root << "\n";
    return 2;
  }
  const int warmups = env_int("BLOCKDAGLU_BENCH_WARMUP", 1);
  const int reps = env_int("BLOCKDAGLU_BENCH_REPETITIONS", 10);
  std::vector<Run> runs = read_baseline(root / "baseline_runs.csv");
  std::string base_alg = runs.empty() ? "" : runs.front().alg;
  std::vector<std::pair<int, double>> xstar_res;
  for (int id : ids) {
    const std::string f = std::to_string(id);
    COOMatrix a;
    int n = 0;
    if (!read_coo(root / f / (f + "_matrix.coo"), n, a)) {
      std::cerr << "case " << id << ": unreadable matrix\n";
      return 2;
    }
    const std::vector<double> b = read_column(root / f / (f + "_rhs.csv"));
    const std::vector<double> xs = read_column(root / f / (f + "_xstar.csv"));
    xstar_res.emplace_back(id, rel_residual(a, n, xs, b));
    for (int w = 0; w < warmups; ++w) (void)run_case(pl, id, 0, a, n, b, xs);
    for (int k = 1; k <= reps; ++k) runs.push_back(run_case(pl, id, k, a, n, b, xs));
    std::cerr << "case " << id << " done (n=" << n << ")\n";
  }

  const char* out_env = std::getenv("BLOCKDAGLU_BENCH_OUT");
  const fs::path out_dir = out_env ? fs::path(out_env) : fs::path();
  if (!out_dir.empty()) fs::create_directories(out_dir);
  auto open = [&](const char* name) -> std::unique_ptr<std::ostream> {
    if (out_dir.empty()) return nullptr;
    return std::make_unique<std::ofstream>(out_dir / name);
  };
  auto fixed = [](std::ostream& o) -> std::ostream& { return o << std::fixed << std::setprecision(3); };
  auto sci = [](std::ostream& o) -> std::ostream& { return o << std::scientific << std::setprecision(6); };

  auto f_runs = open("runs.csv");
  std::ostream& o_runs = f_runs ? *f_runs : std::cout;
  o_runs << "alg,case,repetition,strategy,n,nnz,reorder_ms,lu_ms,solve_ms,total_ms,rel_err,rel_residual,ok\n";
  for (const Run& r : runs) {
    o_runs << r.alg << "," << r.case_id << "," << r.repetition << "," << r.strategy << "," << r.n << "," << r.nnz << ",";
    fixed(o_runs) << r.reorder_ms << "," << r.lu_ms << "," << r.solve_ms << "," << r.total_ms << ",";
    sci(o_runs) << r.rel_err << "," << r.rel_residual << "," << (r.ok ? "1" : "0") << "\n";
  }

  bool precision_pass = true;
  auto f_margin = open("margin.csv");
  std::ostream& o_margin = f_margin ? *f_margin : std::cout;
  o_margin << "case,n,nnz,xstar_rel_residual,ss_rel_residual,ua_rel_residual,eps_suitesparse,err_user,"
              "margin_eps_minus_user,pass_user_le_eps\n";
  for (auto [id, xres] : xstar_res) {
    const Run* ua = median(runs, "BlockDAGLU", id);
    const Run* ss = base_alg.empty() ? nullptr : median(runs, base_alg, id);
    if (!ua) continue;
    const double eps = ss ? ss->rel_err : kInf;
    const bool pass = ua->ok && (!ss || (ss->ok && ua->rel_err <= eps));
    precision_pass = precision_pass && pass;
    o_margin << id << "," << ua->n << "," << ua->nnz << ",";
    sci(o_margin) << xres << "," << (ss ? ss->rel_residual : kInf) << "," << ua->rel_residual << "," << eps << ","
                  << ua->rel_err << "," << (eps - ua->rel_err) << "," << (pass ? "1" : "0") << "\n";
  }

  std::vector<std::string> algs;
  if (!base_alg.empty()) algs.push_back(base_alg);
  algs.push_back("BlockDAGLU");
  auto f_timing = open("timing_summary.csv");
  std::ostream& o_timing = f_timing ? *f_timing : std::cout;
  o_timing << "alg,case,strategy,n,nnz,repetitions,median_total_ms,p95_total_ms,median_rel_err,median_rel_residual,ok\n";
  for (const std::string& alg : algs)
    for (auto [id, xres] : xstar_res) {
      const Run* m = median(runs, alg, id);
      if (!m) continue;
      o_timing << alg << "," << id << "," << m->strategy << "," << m->n << "," << m->nnz << ","
               << runs_of(runs, alg, id).size() << ",";
      fixed(o_timing) << m->total_ms << "," << p95(runs, alg, id) << ",";
      sci(o_timing) << m->rel_err << "," << m->rel_residual << "," << (m->ok ? "1" : "0") << "\n";
    }

  auto f_sum = open("summary.txt");
  std::ostream& o_sum = f_sum ? *f_sum : std::cout;
  double w_base = 0, w_user = 0;
  bool all_ok = true;
  for (const std::string& alg : algs) {
    double w = 0, max_err = 0;
    bool ok = true;
    for (const Run& r : runs)
      if (r.alg == alg) {
        ok = ok && r.ok;
        max_err = std::max(max_err, r.rel_err);
      }
    for (auto [id, xres] : xstar_res) {
      const Run* m = median(runs, alg, id);
      if (!m) ok = false;
      else w += weight_for_n(m->n) * m->total_ms;
    }
    all_ok = all_ok && ok;
    (alg == "BlockDAGLU" ? w_user : w_base) = w;
    o_sum << alg << "_weighted_total_ms,";
    fixed(o_sum) << w << "\n" << alg << "_max_rel_err,";
    sci(o_sum) << max_err << "\n" << alg << "_all_ok," << (ok ? "1" : "0") << "\n";
  }
  if (w_base > 0 && w_user > 0) {
    o_sum << "speedup_suitesparse_over_blockdaglu,";
    fixed(o_sum) << (w_base / w_user) << "\n";
  }
  if (!out_dir.empty()) std::cerr << "wrote benchmark outputs to " << out_dir << "\n";
  if (!all_ok || !precision_pass) {
    std::cerr << "benchmark failed correctness or precision gates\n";
    return 1;
  }
  return 0;
}
