// ep_place: the global-placement part of the reference's driver
// (main/ePlace_main.cpp:35-223) over libeplace_amd.so.
//
//   ep_place -aux <design.aux> [-targetDensity 1.0] [-targetOverflow 0.10]
//            [-loadpl <file.pl>] [-noQP] [-IPiteCount <n>] [-bktrk] [-bb] [-learned <params.npy>] [-LSE]
//            [-bins <n>] [-nomGP] [-internalLegal] [-internalDP] [-outputPath <dir>]
//            [-logCSV <file>]
//            [-dumpDesign <file>]
//
// Reads the Bookshelf design, runs QP (QPPlacer::quadraticPlacement from the
// core centre, ePlace_main.cpp:121-126; -noQP keeps the .pl or -loadpl
// positions), mGP with EplaceNesterovOpt and writes
// <outputPath>/<name>/<name>-eGP.pl (lower-left corners, as
// PlaceDB::outputBookShelf does, placedb.cpp:1022-1060).  Designs with macros
// then run mLG (SAMacroLegalizer), FILLERONLY and cGP (ePlace_main.cpp:174-219)
// unless -nomLG / -nocGP.  Logging goes to stdout.
// -internalLegal runs the Abacus legaliser on the result (reference
// semantics, ePlace_main.cpp:248-259), checks it and writes
// <name>-legal.pl beside the global placement.
// -internalDP then runs the detailed placer on the legal result
// (DetailedPlacer::detailedPlacement, ePlace_main.cpp:271-283), checks it
// and writes <name>-eDP.pl.
// -logCSV writes one row per Nesterov iteration of every stage (stage,
// iteration, overflow, lambda, HPWL, step, 1/gamma_x, evaluations and
// backtracks so far), the fields the reference prints per iteration.
// -dumpDesign writes the parsed arrays (ep_design field order, raw
// little-endian) to a file and exits before any device work.
// -learned runs every stage's Nesterov loop as backtracking led by the
// learned step / penalty controller (the parameters eplace_amd saves, a
// float32 .npy file; no counterpart in the reference).
// -poissonMode selects the transforms' arithmetic (ep_params.poisson_mode:
// 0 FP32 MFMA, the default; 2 split-FP16, the fast mode).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "bookshelf.hpp"
#include "eplacer.hpp"

using namespace eplace_amd;

namespace {

struct Args {
  std::string aux, loadpl, out = ".", dump, log_csv, learned;
  float td = 1.0f, tof = 0.10f;
  bool bktrk = false, bb = false, lse = false, no_mgp = false, legal = false, dp = false, no_qp = false, no_mlg = false, no_cgp = false;
  int bins = 0, qp_iters = 20, poisson_mode = 0;
  // -bb wins over -bktrk, as in EplaceNesterovOpt::opt_step (nesterov.hpp:157-165);
  // -learned (backtracking led by the controller) over both
  int method() const { return !learned.empty() ? STEP_LEARNED : bb ? STEP_BB : bktrk ? STEP_BKTRK : STEP_VANILLA; }
};

int usage(const char *prog) {
  fprintf(stderr,
          "usage: %s -aux <design.aux> [-targetDensity d] [-targetOverflow t] [-loadpl f.pl]\n"
          "          [-noQP] [-IPiteCount n] [-bktrk] [-bb] [-LSE] [-bins n] [-nomGP] [-nomLG] [-nocGP]\n"
          "          [-learned params.npy] [-poissonMode m] [-internalLegal] [-internalDP] [-logCSV file]\n"
          "          [-outputPath dir] [-dumpDesign file]\n",
          prog);
  return 2;
}

bool parse(int argc, char **argv, Args &a) {
  for (int i = 1; i < argc; ++i) {
    std::string k = argv[i];
    auto val = [&](std::string &dst) {
      if (i + 1 >= argc) return false;
      dst = argv[++i];
      return true;
    };
    std::string v;
    if (k == "-aux") {
      if (!val(a.aux)) return false;
    } else if (k == "-loadpl") {
      if (!val(a.loadpl)) return false;
    } else if (k == "-outputPath") {
      if (!val(a.out)) return false;
    } else if (k == "-dumpDesign") {
      if (!val(a.dump)) return false;
    } else if (k == "-targetDensity") {
      if (!val(v)) return false;
      a.td = (float)atof(v.c_str());
    } else if (k == "-targetOverflow") {
      if (!val(v)) return false;
      a.tof = (float)atof(v.c_str());
    } else if (k == "-bins") {
      if (!val(v)) return false;
      a.bins = atoi(v.c_str());
    } else if (k == "-bktrk") {
      a.bktrk = true;
    } else if (k == "-bb") {
      a.bb = true;
    } else if (k == "-LSE") {
      a.lse = true;
    } else if (k == "-nomGP") {
      a.no_mgp = true;
    } else if (k == "-internalLegal") {
      a.legal = true;
    } else if (k == "-internalDP") {
      a.dp = true;
    } else if (k == "-logCSV") {
      if (!val(a.log_csv)) return false;
    } else if (k == "-learned") {
      if (!val(a.learned)) return false;
    } else if (k == "-poissonMode") {
      if (!val(v)) return false;
      a.poisson_mode = atoi(v.c_str());
    } else if (k == "-IPiteCount") {
      if (!val(v)) return false;
      a.qp_iters = atoi(v.c_str());
    } else if (k == "-noQP") {
      a.no_qp = true;
    } else if (k == "-nomLG") {
      a.no_mlg = true;
    } else if (k == "-nocGP") {
      a.no_cgp = true;
    } else if (k == "-noLegal") {
      // these stages are not run here; accepted for command-line compatibility
    } else {
      fprintf(stderr, "unknown option %s\n", k.c_str());
      return false;
    }
  }
  return !a.aux.empty();
}

void mkdirs(const std::string &path) {
  std::string cur;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || path[i] == '/') {
      if (!cur.empty()) mkdir(cur.c_str(), 0755);
    }
    if (i < path.size()) cur += path[i];
  }
}

template <typename T>
void put(FILE *f, const std::vector<T> &v) {
  int64_t n = (int64_t)v.size();
  fwrite(&n, sizeof n, 1, f);
  if (n) fwrite(v.data(), sizeof(T), v.size(), f);
}

void dump_design(const BookshelfDesign &d, const std::string &path) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  put(f, d.node_w), put(f, d.node_h), put(f, d.node_x), put(f, d.node_y);
  put(f, d.term_w), put(f, d.term_h), put(f, d.term_x), put(f, d.term_y), put(f, d.term_ni);
  put(f, d.net_start), put(f, d.pin_obj), put(f, d.pin_dx), put(f, d.pin_dy);
  put(f, d.row_bottom), put(f, d.row_height), put(f, d.row_x0), put(f, d.row_x1);
  fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
  Args a;
  if (!parse(argc, argv, a)) return usage(argv[0]);
  try {
    // read and checked before any device work
    const std::vector<float> ctrl_params = a.learned.empty() ? std::vector<float>() : Controller::load(a.learned);
    BookshelfDesign d = read_bookshelf(a.aux);
    if (!a.loadpl.empty()) load_pl(d, a.loadpl);
    printf("design %s: %zu nodes, %zu terminals, %zu nets, %zu pins, %zu rows\n", d.name.c_str(),
           d.node_w.size(), d.term_w.size(), d.net_start.size() - 1, d.pin_obj.size(), d.row_bottom.size());
    if (!a.dump.empty()) {
      dump_design(d, a.dump);
      return 0;
    }
    ep_design view = d.view();
    if (!a.no_qp) {
      // QPPlacer::quadraticPlacement (qplace.cpp:20-103)
      std::vector<float> xy(2 * d.node_x.size());
      for (size_t i = 0; i < d.node_x.size(); ++i) xy[2 * i] = d.node_x[i], xy[2 * i + 1] = d.node_y[i];
      std::vector<double> log;
      ep_qp_stats qs = QPPlacer(view).quadraticPlacement(xy, a.qp_iters, &log);
      printf("INFO:  initial HPWL: %.6lf\n", qs.initial_hpwl);
      printf("INFO:  The matrix size is %zu\n", d.node_x.size());
      for (int i = 0; i < qs.iterations; ++i)
        printf("INFO:  at iteration number %3d,  CG Error %.6lf,  HPWL %.6lf\n", i,
               std::max(log[4 * i + 1], log[4 * i + 2]), log[4 * i]);
      printf("QP time: %.4f s\n", qs.seconds);
      for (size_t i = 0; i < d.node_x.size(); ++i) d.node_x[i] = xy[2 * i], d.node_y[i] = xy[2 * i + 1];
    } else {
      printf("QP skipped (-noQP): starting from the .pl positions\n");
    }
    PlacerOptions o;
    o.bins_x = o.bins_y = a.bins;
    o.lse = a.lse;
    o.poisson_mode = a.poisson_mode;
    EPlacer_2D placer(view, a.td, o);
    placer.initialization();
    // -learned: the controller's saved parameters, on the placer's device
    std::unique_ptr<Controller> ctrl;
    if (!a.learned.empty()) ctrl.reset(new Controller(ctrl_params, o.device));
    ep_info info = placer.info();
    printf("bins %dx%d, %lld fillers, target density %.3f\n", info.bins_x, info.bins_y, (long long)info.num_fillers, a.td);

    FILE *csv = nullptr;
    if (!a.log_csv.empty()) {
      csv = fopen(a.log_csv.c_str(), "w");
      if (!csv) throw std::runtime_error("cannot write " + a.log_csv);
      fprintf(csv, "stage,iteration,overflow,lambda,hpwl,step,inv_gamma_x,evaluations,backtracks\n");
    }
    // one CSV row per iteration of a stage's run (EP_LOG_* columns)
    auto log_rows = [&](const char *stage, const EplaceNesterovOpt<VECTOR_3D> &o) {
      if (!csv) return;
      const std::vector<double> &g = o.log();
      for (size_t r = 0; r + EP_LOG_COLS <= g.size(); r += EP_LOG_COLS)
        fprintf(csv, "%s,%d,%.9g,%.9g,%.17g,%.9g,%.9g,%d,%d\n", stage, (int)g[r + EP_LOG_ITER], g[r + EP_LOG_TAU],
                g[r + EP_LOG_LAMBDA], g[r + EP_LOG_HPWL], g[r + EP_LOG_STEP], g[r + EP_LOG_INV_GAMMA],
                (int)g[r + EP_LOG_EVALS], (int)g[r + EP_LOG_BACKTRACKS]);
    };
    if (!a.no_mgp) {
      EplaceNesterovOpt<VECTOR_3D> opt(&placer, a.method(), a.tof);
      opt.setController(ctrl.get());
      opt.keepLog(csv != nullptr);
      opt.opt();
      log_rows("mGP", opt);
      const ep_run_stats &s = opt.last();
      printf("mGP finished: %d iterations, %d evaluations, overflow %.4f\n", s.iterations, s.evaluations,
             s.final_overflow);
      printf("Final HPWL: %.6e\n", s.final_hpwl);
      printf("mGP time: %.4f s\n", s.seconds);
    }
    size_t macros = 0;
    for (float h : d.node_h)  // PlaceDB::addNode (placedb.cpp:102-111)
      macros += (double)(h - (float)view.common_row_height) > 1e-15;
    if (macros > 0 && !a.no_mlg) {
      // ePlace_main.cpp:174-219
      SAMacroLegalizer mlg(&placer, view);
      mlg.setTargetDensity(a.td);
      ep_mlg_stats ms = mlg.legalization();
      printf("mLG finished: %d macros, %d rounds, %lld moves, overlap %d; HPWL after mLG: %.6e (%.3f s)\n", ms.macros,
             ms.outer_iterations, (long long)ms.perturbations, ms.final_overlap, ms.final_hpwl, ms.seconds);
      if (!a.no_cgp) {
        placer.switch2FillerOnly();
        EplaceNesterovOpt<VECTOR_3D> fo(&placer, a.method(), a.tof);
        fo.setController(ctrl.get());
        fo.keepLog(csv != nullptr);
        fo.opt();
        log_rows("FILLERONLY", fo);
        printf("filler placement finished: %d iterations\n", fo.last().iterations);
        placer.switch2cGP();
        EplaceNesterovOpt<VECTOR_3D> cg(&placer, a.method(), a.tof);
        cg.setController(ctrl.get());
        cg.keepLog(csv != nullptr);
        cg.opt();
        log_rows("cGP", cg);
        const ep_run_stats &s = cg.last();
        printf("cGP finished: %d iterations, overflow %.4f\n", s.iterations, s.final_overflow);
        printf("cGP Final HPWL: %.6e\n", s.final_hpwl);
      }
    } else if (macros > 0) {
      printf("%zu macros: mLG skipped (-nomLG)\n", macros);
    }

    std::vector<float> xy(2 * d.node_w.size());
    std::vector<float> all(2 * (size_t)info.num_objects);
    placer.check(ep_read(placer.context(), EP_ARR_POSITIONS, all.data(), (int64_t)all.size()));
    std::copy(all.begin(), all.begin() + xy.size(), xy.begin());
    if (csv) {
      fclose(csv);
      printf("wrote %s\n", a.log_csv.c_str());
    }
    std::string dir = a.out + "/" + d.name + "/";
    mkdirs(dir);
    std::string pl = dir + d.name + "-eGP.pl";
    write_pl(d, xy, pl);
    printf("wrote %s\n", pl.c_str());
    if (a.legal) {
      std::vector<float> lxy(xy.size());
      ep_legal_params lp{};
      lp.site_step = d.site_width;
      lp.mode = 0;
      ep_legal_stats ls{};
      if (ep_legalize_abacus(&view, xy.data(), &lp, lxy.data(), &ls) != EP_OK)
        throw std::runtime_error(std::string("legalisation: ") + ep_host_last_error());
      ep_legal_check chk{};
      if (ep_check_legal(&view, lxy.data(), d.site_width, &chk) != EP_OK)
        throw std::runtime_error(std::string("legality check: ") + ep_host_last_error());
      const bool ok = chk.overlaps == 0 && chk.off_row == 0 && chk.off_site == 0 && chk.out_of_core == 0 &&
                      chk.in_blockage == 0 && chk.macro_overlaps == 0;
      printf("legalised %lld cells: total displacement %.1f, max %.2f; %s (overlaps %lld, off-row %lld, "
             "off-site %lld, in blockage %lld)\n",
             (long long)ls.cells, ls.total_displacement, ls.max_displacement, ok ? "legal" : "NOT legal",
             (long long)chk.overlaps, (long long)chk.off_row, (long long)chk.off_site, (long long)chk.in_blockage);
      std::string lpl = dir + d.name + "-legal.pl";
      write_pl(d, lxy, lpl);
      printf("wrote %s\n", lpl.c_str());
      if (!ok) return 3;
      if (a.dp) {
        std::vector<float> dxy(xy.size());
        ep_dp_params dpp{};
        dpp.site_step = d.site_width;
        dpp.passes = 2;
        ep_dp_stats ds{};
        if (ep_detailed_place(&view, lxy.data(), &dpp, dxy.data(), &ds) != EP_OK)
          throw std::runtime_error(std::string("detailed placement: ") + ep_host_last_error());
        ep_legal_check dchk{};
        if (ep_check_legal(&view, dxy.data(), d.site_width, &dchk) != EP_OK)
          throw std::runtime_error(std::string("legality check: ") + ep_host_last_error());
        const bool dok = dchk.overlaps == 0 && dchk.off_row == 0 && dchk.off_site == 0 && dchk.out_of_core == 0 &&
                         dchk.in_blockage == 0 && dchk.macro_overlaps == 0;
        printf("HPWL before detailed placement: %.0f\nHPWL after detailed placement: %.0f (%.2fs); %s\n",
               ds.hpwl_before, ds.hpwl_after, ds.seconds, dok ? "legal" : "NOT legal");
        std::string dpl = dir + d.name + "-eDP.pl";
        write_pl(d, dxy, dpl);
        printf("wrote %s\n", dpl.c_str());
        if (!dok) return 3;
      }
    } else if (a.dp) {
      printf("-internalDP runs after -internalLegal only; skipped\n");
    }
  } catch (const std::exception &e) {
    fprintf(stderr, "ep_place: %s\n", e.what());
    return 1;
  }
  return 0;
}
