// Golden-vector driver for the reference ePlace placer (test infrastructure).
//
// This file is ours; it is linked against the reference's own, unmodified
// sources (EPlace/, PlaceDB/, PlaceCommon/, Parser/, FFT/, QPlace/, Plot/,
// Legalization/, DetailedPlacement/) by oracle/Makefile.ref and built only
// into oracle/_ref/.  It drives the reference's public call surface the way
// main/ePlace_main.cpp:35-284 does, and dumps the state the parity tests
// compare against as .npy files plus a text file of scalars.
//
// Modes
//   gp   -aux F.aux -targetDensity TD [-targetOverflow T] [-bktrk] [-LSE]
//        [-noQP] [-loadpl F.pl] [-iters K] [-dumpEvery M] [-mixed] [-legal]
//        [-bins N] -out DIR
//        (-mixed: for designs with macros, mLG, FILLERONLY and cGP after mGP;
//        -bins: an N x N bin grid instead of the reference's bin rule)
//   fft  -nx NX -ny NY -bx BX -by BY -in RHO.f32 -out DIR
//   time -aux F.aux -targetDensity TD [-noQP] [-loadpl F.pl] [-iters K]
//        (wall time of K vanilla iterations after initialisation, 1 thread)
//   legal -aux F.aux [-centers C.f32] -out DIR
//        (AbacusLegalizer on the node centres in C.f32, float pairs in
//        dbNodes order set through setModuleCenter_2D; default: the .pl)
//   qp   -aux F.aux -noQP [-loadpl F.pl] [-center | -centers C.f32] [-IPiteCount K] [-timeOnly] -out DIR
//        (QPPlacer::createSparseMatrix at the loaded positions, or at the
//        core centre with -center, dumped as summed COO plus b and the
//        guess; then QPPlacer::quadraticPlacement, whose per-iteration
//        HPWL and solver error go to stdout, and the node centres after it;
//        -timeOnly: only the wall time and HPWL of quadraticPlacement)
//   mlg  -aux F.aux -noQP [-centers C.f32] [-targetDensity TD] [-randSkip K] -out DIR
//        (SAMacroLegalizer on the node centres after K rand() draws; the
//        centres after it, its final cost terms and the next rand() value)
//   dp   -aux F.aux -noQP [-centers C.f32] [-passes P] -out DIR
//        (AbacusLegalizer on the centres, then DetailedPlacer's rounds as
//        detailedPlacement runs them, the positions after each operator of
//        each round: lr<i>, ism<i>, gs<i>, with their HPWL)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>
#include <map>
#include <fstream>
#include <sstream>
#include <iostream>
#include <sys/stat.h>
#include <sys/time.h>

#include "global.h"
#include "objects.h"
#include "placedb.h"
#include "parser.h"
#include "fft.h"
#include "eplace.h"
#include "qplace.h"
#include "legalizer.h"
#include "plot.h"
#include "opt.hpp"
#include "optutils.hpp"
// The optimiser keeps its state and step functions private; the harness
// needs to single-step it and read that state, so widen access for this one
// header only.
#define private public
#include "nesterov.hpp"
#include "detailed.h"
#undef private

namespace {

std::string g_out;

void write_npy(const std::string &name, const char *descr, const void *data,
               size_t elem, const std::vector<size_t> &shape) {
  std::string hdr = "{'descr': '";
  hdr += descr;
  hdr += "', 'fortran_order': False, 'shape': (";
  size_t n = 1;
  for (size_t i = 0; i < shape.size(); ++i) {
    hdr += std::to_string(shape[i]);
    hdr += (shape.size() == 1 || i + 1 < shape.size()) ? ", " : "";
    n *= shape[i];
  }
  hdr += "), }";
  size_t total = 10 + hdr.size() + 1;
  size_t pad = (64 - total % 64) % 64;
  hdr.append(pad, ' ');
  hdr += '\n';
  std::string path = g_out + "/" + name + ".npy";
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) { perror(path.c_str()); exit(2); }
  const char magic[] = "\x93NUMPY\x01\x00";
  fwrite(magic, 1, 8, f);
  uint16_t hl = (uint16_t)hdr.size();
  fwrite(&hl, 2, 1, f);
  fwrite(hdr.data(), 1, hdr.size(), f);
  fwrite(data, elem, n, f);
  fclose(f);
}

void npy_f32(const std::string &n, const std::vector<float> &v, size_t cols = 0) {
  if (cols) write_npy(n, "<f4", v.data(), 4, {v.size() / cols, cols});
  else write_npy(n, "<f4", v.data(), 4, {v.size()});
}
void npy_f64(const std::string &n, const std::vector<double> &v, size_t cols = 0) {
  if (cols) write_npy(n, "<f8", v.data(), 8, {v.size() / cols, cols});
  else write_npy(n, "<f8", v.data(), 8, {v.size()});
}
void npy_i32(const std::string &n, const std::vector<int32_t> &v) {
  write_npy(n, "<i4", v.data(), 4, {v.size()});
}

std::ofstream g_scalars;
template <class T>
void scalar(const std::string &k, T v) {
  g_scalars.precision(17);
  g_scalars << k << " " << v << "\n";
}

std::string arg_str(int argc, char **argv, const char *key, const char *def) {
  for (int i = 1; i + 1 < argc; ++i)
    if (argv[i][0] == '-' && strcmp(argv[i] + 1, key) == 0) return argv[i + 1];
  return def;
}
bool arg_flag(int argc, char **argv, const char *key) {
  for (int i = 1; i < argc; ++i)
    if (argv[i][0] == '-' && strcmp(argv[i] + 1, key) == 0) return true;
  return false;
}

double now() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

void load_design(int argc, char **argv, BookshelfParser &parser, PlaceDB *db) {
  std::string aux = arg_str(argc, argv, "aux", "");
  size_t pos = aux.rfind('/');
  gArg.Override("path", pos == std::string::npos ? std::string("./") : aux.substr(0, pos + 1));
  parser.ReadFile(aux, *db);
  std::string pl = arg_str(argc, argv, "loadpl", "");
  if (!pl.empty()) parser.ReadPLFile(pl, *db, false);
  if (!arg_flag(argc, argv, "noQP")) {
    QPPlacer qp(db);
    qp.quadraticPlacement();
  }
}

// Design as the reference holds it after parsing (and QP, if run).
void dump_design(PlaceDB *db) {
  std::vector<float> nodes, terms;
  std::vector<int32_t> nodeMacro, termNI;
  for (Module *m : db->dbNodes) {
    nodes.insert(nodes.end(), {m->getWidth(), m->getHeight(), m->getCenter().x, m->getCenter().y});
    nodeMacro.push_back(m->isMacro ? 1 : 0);
  }
  for (Module *m : db->dbTerminals) {
    terms.insert(terms.end(), {m->getWidth(), m->getHeight(), m->getCenter().x, m->getCenter().y,
                               m->getLocation().x, m->getLocation().y});
    termNI.push_back(m->isNI ? 1 : 0);
  }
  npy_f32("nodes_whxy", nodes, 4);
  npy_i32("nodes_macro", nodeMacro);
  if (!terms.empty()) npy_f32("terms_whxyll", terms, 6);
  npy_i32("terms_ni", termNI);
  // Pins in dbPins order: the parser adds them net by net (parser.cpp:473-680).
  std::map<Module *, int32_t> objIdx;
  for (size_t i = 0; i < db->dbNodes.size(); ++i) objIdx[db->dbNodes[i]] = (int32_t)i;
  for (size_t i = 0; i < db->dbTerminals.size(); ++i)
    objIdx[db->dbTerminals[i]] = (int32_t)(db->dbNodes.size() + i);
  std::vector<int32_t> pinObj, netStart;
  std::vector<float> pinOff;
  for (Net *n : db->dbNets) {
    netStart.push_back((int32_t)pinObj.size());
    for (Pin *p : n->netPins) {
      pinObj.push_back(objIdx[p->module]);
      pinOff.push_back(p->offset.x);
      pinOff.push_back(p->offset.y);
    }
  }
  netStart.push_back((int32_t)pinObj.size());
  npy_i32("pin_obj", pinObj);
  npy_f32("pin_offset", pinOff, 2);
  npy_i32("net_start", netStart);
  std::vector<double> rows;
  for (SiteRow &r : db->dbSiteRows)
    rows.insert(rows.end(), {r.bottom, r.height, (double)r.start.x, (double)r.end.x, r.step});
  npy_f64("rows", rows, 5);
  scalar("core_llx", db->coreRegion.ll.x);
  scalar("core_lly", db->coreRegion.ll.y);
  scalar("core_urx", db->coreRegion.ur.x);
  scalar("core_ury", db->coreRegion.ur.y);
  scalar("total_row_area", db->totalRowArea);
  scalar("common_row_height", db->commonRowHeight);
  scalar("hpwl_design", db->calcHPWL());
}

std::vector<float> positions(const std::vector<Module *> &objs) {
  std::vector<float> v;
  v.reserve(objs.size() * 2);
  for (Module *m : objs) {
    v.push_back(m->getCenter().x);
    v.push_back(m->getCenter().y);
  }
  return v;
}

std::vector<float> vec3_xy(const std::vector<VECTOR_3D> &g) {
  std::vector<float> v;
  v.reserve(g.size() * 2);
  for (const VECTOR_3D &e : g) {
    v.push_back(e.x);
    v.push_back(e.y);
  }
  return v;
}

// Evaluator state after one totalGradientUpdate (eplace.cpp:687-738).
void dump_eval(EPlacer_2D *ep, const std::string &tag) {
  int nx = ep->binDimension.x, ny = ep->binDimension.y;
  std::vector<float> node(nx * ny), filler(nx * ny), term(nx * ny), base(nx * ny),
      ex(nx * ny), ey(nx * ny), phi(nx * ny);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      Bin_2D *b = ep->bins[i][j];
      size_t k = (size_t)i * ny + j;
      node[k] = b->nodeDensity;
      filler[k] = b->fillerDensity;
      term[k] = b->terminalDensity;
      base[k] = b->baseDensity;
      ex[k] = b->E.x;
      ey[k] = b->E.y;
      phi[k] = b->phi;
    }
  npy_f32(tag + "map_node", node, ny);
  npy_f32(tag + "map_filler", filler, ny);
  npy_f32(tag + "map_terminal", term, ny);
  npy_f32(tag + "map_base", base, ny);
  npy_f32(tag + "map_ex", ex, ny);
  npy_f32(tag + "map_ey", ey, ny);
  npy_f32(tag + "map_phi", phi, ny);
  npy_f32(tag + "grad_density", vec3_xy(ep->densityGradient), 2);
  npy_f32(tag + "grad_wirelength", vec3_xy(ep->wirelengthGradient), 2);
  npy_f32(tag + "grad_total", vec3_xy(ep->totalGradient), 2);
  npy_f32(tag + "positions", positions(ep->ePlaceNodesAndFillers), 2);
  scalar(tag + "overflow", ep->globalDensityOverflow);
  scalar(tag + "inv_gamma_x", ep->invertedGamma.x);
  scalar(tag + "inv_gamma_y", ep->invertedGamma.y);
  scalar(tag + "lambda", ep->lambda);
  scalar(tag + "last_hpwl", ep->lastHPWL);
}

// EPlacer_2D::binInitialization (eplace.cpp:221-412) for a given grid: the
// reference derives the grid from the average node area and caps it at
// 1024; the workload configurations name their grids (16x16 for 100 cells,
// 1024x1024 and 2048x2048), so the harness builds the bins, the terminal
// density and the base density for nx x ny bins the way that function
// builds them for its own grid.
void bin_initialization_at(EPlacer_2D *ep, PlaceDB *db, int nx, int ny) {
  const float cw = db->coreRegion.getWidth(), ch = db->coreRegion.getHeight();
  ep->binDimension.x = nx;
  ep->binDimension.y = ny;
  ep->binStep.x = float_div(cw, nx);
  ep->binStep.y = float_div(ch, ny);
  const float sx = ep->binStep.x, sy = ep->binStep.y;
  ep->bins.assign(nx, std::vector<Bin_2D *>());
  for (int i = 0; i < nx; i++) {
    ep->bins[i].resize(ny);
    for (int j = 0; j < ny; j++) {
      Bin_2D *b = new Bin_2D();
      b->ll.x = i * sx + db->coreRegion.ll.x;
      b->ll.y = j * sy + db->coreRegion.ll.y;
      b->width = sx;
      b->height = sy;
      b->ur.x = b->ll.x + b->width;
      b->ur.y = b->ll.y + b->height;
      b->area = sx * sy;
      b->center.x = b->ll.x + (float)0.5 * b->width;
      b->center.y = b->ll.y + (float)0.5 * b->height;
      ep->bins[i][j] = b;
    }
  }
  const float td = ep->targetDensity;
  for (Module *t : db->dbTerminals) {
    const POS_2D ll = t->getLL_2D(), ur = t->getUR_2D();
    if (ll.x < db->coreRegion.ll.x || ll.x + t->getWidth() > db->coreRegion.ur.x) continue;
    if (ll.y < db->coreRegion.ll.y || ll.y + t->getHeight() > db->coreRegion.ur.y) continue;
    const int x0 = (int)((ll.x - db->coreRegion.ll.x) / sx), y0 = (int)((ll.y - db->coreRegion.ll.y) / sy);
    const int x1 = std::min((int)((ur.x - db->coreRegion.ll.x) / sx), nx - 1);
    const int y1 = std::min((int)((ur.y - db->coreRegion.ll.y) / sy), ny - 1);
    for (int i = x0; i <= x1; i++)
      for (int j = y0; j <= y1; j++)
        ep->bins[i][j]->terminalDensity += td * getOverlapArea_2D(ep->bins[i][j]->ll, ep->bins[i][j]->ur, ll, ur);
  }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      Bin_2D *b = ep->bins[i][j];
      float avail = 0;
      for (SiteRow r : db->dbSiteRows) avail += getOverlapArea_2D(b->ll, b->ur, r.getLL_2D(), r.getUR_2D());
      b->baseDensity = float_equal(b->area, avail) ? 0 : td * (b->area - avail);
    }
}

int mode_gp(int argc, char **argv) {
  BookshelfParser parser;
  PlaceDB *db = new PlaceDB();
  load_design(argc, argv, parser, db);
  dump_design(db);

  float td = (float)atof(arg_str(argc, argv, "targetDensity", "1.0").c_str());
  EPlacer_2D *ep = new EPlacer_2D(db);
  ep->setTargetDensity(td);
  // initialization() = fillers, bins, gradient vectors, first evaluation with
  // lambda = 0, then lambda0 (eplace.cpp:39-50).  Run the parts one by one so
  // the first evaluation can be dumped before lambda0 is set.
  ep->fillerInitialization();
  const int fixedBins = atoi(arg_str(argc, argv, "bins", "0").c_str());
  if (fixedBins > 0) bin_initialization_at(ep, db, fixedBins, fixedBins);
  else ep->binInitialization();
  ep->gradientVectorInitialization();
  npy_f32("fillers_init", positions(ep->ePlaceFillers), 2);
  std::vector<float> fw;
  for (Module *f : ep->ePlaceFillers) { fw.push_back(f->getWidth()); fw.push_back(f->getHeight()); }
  npy_f32("fillers_wh", fw, 2);
  scalar("num_fillers", ep->ePlaceFillers.size());
  scalar("bins_x", ep->binDimension.x);
  scalar("bins_y", ep->binDimension.y);
  scalar("bin_step_x", ep->binStep.x);
  scalar("bin_step_y", ep->binStep.y);
  scalar("std_cell_area", ep->ePlaceStdCellArea);
  scalar("macro_area", ep->ePlaceMacroArea);
  // Lower-left corners as stored (Module::coor); the scatter reads these, and
  // for objects placed by lower-left they need not equal centre - w/2.
  std::vector<float> ll;
  for (Module *m : ep->ePlaceNodesAndFillers) {
    ll.push_back(m->getLL_2D().x);
    ll.push_back(m->getLL_2D().y);
  }
  npy_f32("init_ll", ll, 2);
  ep->totalGradientUpdate();
  dump_eval(ep, "init_");
  ep->penaltyFactorInitilization();
  scalar("lambda0", ep->lambda);
  scalar("hpwl0", ep->lastHPWL);

  int iters = atoi(arg_str(argc, argv, "iters", "1000000").c_str());
  int dumpEvery = atoi(arg_str(argc, argv, "dumpEvery", "0").c_str());
  EplaceNesterovOpt<VECTOR_3D> opt(ep);
  opt.init();
  std::vector<double> traj;  // per iteration: k, overflow, lambda, hpwl, invgx, invgy, a_k
  double t0 = now();
  int k = 0;
  while (!opt.stop_condition() && k < iters) {
    opt.opt_step();
    traj.insert(traj.end(), {(double)k, ep->globalDensityOverflow, ep->lambda, ep->lastHPWL,
                             ep->invertedGamma.x, ep->invertedGamma.y, opt.NS_opt_param});
    if (k < 3 || (dumpEvery > 0 && k % dumpEvery == 0)) {
      std::string tag = "it" + std::to_string(k) + "_";
      npy_f32(tag + "main", vec3_xy(opt.cur_iter.main_solution), 2);
      npy_f32(tag + "ref_unclamped", vec3_xy(opt.cur_iter.reference_solution), 2);
      npy_f32(tag + "grad_used", vec3_xy(opt.last_iter.gradient), 2);
      npy_f32(tag + "ref_used", vec3_xy(opt.last_iter.reference_solution), 2);
      npy_f32(tag + "db_positions", positions(ep->ePlaceNodesAndFillers), 2);
    }
    ++k;
  }
  double gpTime = now() - t0;
  bool stopped = opt.stop_condition();
  npy_f64("trajectory", traj, 7);
  scalar("iterations", k);
  scalar("stopped", stopped ? 1 : 0);
  scalar("gp_time", gpTime);
  scalar("final_overflow", ep->globalDensityOverflow);
  opt.wrap_up();
  npy_f32("final_positions", positions(ep->ePlaceNodesAndFillers), 2);
  scalar("final_hpwl", db->calcHPWL());
  if (arg_flag(argc, argv, "mixed") && db->dbMacroCount > 0) {
    // ePlace_main.cpp:174-219: mLG, then FILLERONLY and cGP with the same
    // optimiser object
    SAMacroLegalizer mlg(db);
    mlg.setTargetDensity(td);
    mlg.legalization();
    npy_f32("mixed_mlg_positions", positions(db->dbNodes), 2);
    scalar("mixed_mlg_hpwl", db->calcHPWL());
    scalar("mixed_mgp_iteration_count", ep->mGPIterationCount);
    ep->switch2FillerOnly();
    npy_f32("mixed_fo_fillers", positions(ep->ePlaceFillers), 2);
    opt.opt();
    npy_f32("mixed_fo_positions", positions(ep->ePlaceNodesAndFillers), 2);
    scalar("mixed_fo_iterations", opt.iter_count);
    scalar("mixed_fo_overflow", ep->globalDensityOverflow);
    scalar("mixed_fo_lambda", ep->lambda);
    ep->switch2cGP();
    scalar("mixed_cgp_lambda_start", ep->lambda);
    opt.opt();
    npy_f32("mixed_cgp_positions", positions(ep->ePlaceNodesAndFillers), 2);
    scalar("mixed_cgp_iterations", opt.iter_count);
    scalar("mixed_cgp_overflow", ep->globalDensityOverflow);
    scalar("mixed_cgp_lambda", ep->lambda);
    scalar("mixed_cgp_hpwl", db->calcHPWL());
  }
  if (arg_flag(argc, argv, "legal")) {
    AbacusLegalizer legal(db);
    legal.legalization();
    npy_f32("legal_positions", positions(db->dbNodes), 2);
    scalar("legal_hpwl", db->calcHPWL());
  }
  return 0;
}

int mode_fft(int argc, char **argv) {
  int nx = atoi(arg_str(argc, argv, "nx", "0").c_str());
  int ny = atoi(arg_str(argc, argv, "ny", "0").c_str());
  float bx = (float)atof(arg_str(argc, argv, "bx", "1").c_str());
  float by = (float)atof(arg_str(argc, argv, "by", "1").c_str());
  std::string in = arg_str(argc, argv, "in", "");
  std::vector<float> rho((size_t)nx * ny);
  FILE *f = fopen(in.c_str(), "rb");
  if (!f || fread(rho.data(), 4, rho.size(), f) != rho.size()) { fprintf(stderr, "bad input\n"); return 2; }
  fclose(f);
  replace::FFT_2D fft(nx, ny, bx, by);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) fft.updateDensity(i, j, rho[(size_t)i * ny + j]);
  fft.doFFT();
  std::vector<float> ex(rho.size()), ey(rho.size()), phi(rho.size());
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      auto e = fft.getElectroForce(i, j);
      ex[(size_t)i * ny + j] = e.first;
      ey[(size_t)i * ny + j] = e.second;
      phi[(size_t)i * ny + j] = fft.getElectroPhi(i, j);
    }
  npy_f32("ex", ex, ny);
  npy_f32("ey", ey, ny);
  npy_f32("phi", phi, ny);
  return 0;
}

// Wall time of the reference's per-iteration path on one core.
int mode_time(int argc, char **argv) {
  BookshelfParser parser;
  PlaceDB *db = new PlaceDB();
  double t0 = now();
  load_design(argc, argv, parser, db);
  float td = (float)atof(arg_str(argc, argv, "targetDensity", "1.0").c_str());
  EPlacer_2D *ep = new EPlacer_2D(db);
  ep->setTargetDensity(td);
  ep->initialization();
  double tInit = now() - t0;
  int iters = atoi(arg_str(argc, argv, "iters", "3").c_str());
  EplaceNesterovOpt<VECTOR_3D> opt(ep);
  opt.init();
  double t1 = now();
  int k = 0;
  while (k < iters) { opt.opt_step(); ++k; }
  double tIter = now() - t1;
  t1 = now();
  ep->totalGradientUpdate();
  double tGrad = now() - t1;
  scalar("init_time", tInit);
  scalar("iterations", k);
  scalar("iter_time", tIter);
  scalar("grad_time", tGrad);
  fprintf(stderr, "init %.3f s, %d iterations %.3f s (%.4f it/s), one gradient %.3f s\n", tInit, k,
          tIter, k / tIter, tGrad);
  return 0;
}

int mode_legal(int argc, char **argv) {
  BookshelfParser parser;
  PlaceDB *db = new PlaceDB();
  load_design(argc, argv, parser, db);
  dump_design(db);
  std::string cin = arg_str(argc, argv, "centers", "");
  if (!cin.empty()) {
    std::vector<float> xy(db->dbNodes.size() * 2);
    FILE *f = fopen(cin.c_str(), "rb");
    if (!f || fread(xy.data(), 4, xy.size(), f) != xy.size()) {
      fprintf(stderr, "bad centers file\n");
      return 2;
    }
    fclose(f);
    for (size_t i = 0; i < db->dbNodes.size(); ++i) db->setModuleCenter_2D(db->dbNodes[i], xy[2 * i], xy[2 * i + 1]);
  }
  npy_f32("input_positions", positions(db->dbNodes), 2);
  scalar("input_hpwl", db->calcHPWL());
  AbacusLegalizer legal(db);
  legal.legalization();
  npy_f32("legal_positions", positions(db->dbNodes), 2);
  scalar("legal_hpwl", db->calcHPWL());
  return 0;
}

// DetailedPlacer::detailedPlacement (detailed.cpp:74-85) one operator at a
// time, so each operator's result is dumped.
int mode_dp(int argc, char **argv) {
  BookshelfParser parser;
  PlaceDB *db = new PlaceDB();
  load_design(argc, argv, parser, db);
  dump_design(db);
  std::string cin = arg_str(argc, argv, "centers", "");
  if (!cin.empty()) {
    std::vector<float> xy(db->dbNodes.size() * 2);
    FILE *f = fopen(cin.c_str(), "rb");
    if (!f || fread(xy.data(), 4, xy.size(), f) != xy.size()) {
      fprintf(stderr, "bad centers file\n");
      return 2;
    }
    fclose(f);
    for (size_t i = 0; i < db->dbNodes.size(); ++i) db->setModuleCenter_2D(db->dbNodes[i], xy[2 * i], xy[2 * i + 1]);
  }
  AbacusLegalizer legal(db);
  legal.legalization();
  npy_f32("legal_positions", positions(db->dbNodes), 2);
  scalar("legal_hpwl", db->calcHPWL());
  const int passes = atoi(arg_str(argc, argv, "passes", "2").c_str());
  DetailedPlacer dp(db);
  struct timeval t0, t1;
  gettimeofday(&t0, nullptr);
  dp.initialization();
  for (int i = 0; i < passes; ++i) {
    const std::string k = std::to_string(i);
    dp.runLocalReordering();
    npy_f32("lr" + k, positions(db->dbNodes), 2);
    scalar("lr" + k + "_hpwl", db->calcHPWL());
    dp.runISM();
    npy_f32("ism" + k, positions(db->dbNodes), 2);
    scalar("ism" + k + "_hpwl", db->calcHPWL());
    dp.runGlobalSwap();
    npy_f32("gs" + k, positions(db->dbNodes), 2);
    scalar("gs" + k + "_hpwl", db->calcHPWL());
  }
  gettimeofday(&t1, nullptr);
  scalar("dp_hpwl", db->calcHPWL());
  scalar("dp_seconds", (t1.tv_sec - t0.tv_sec) + 1e-6 * (t1.tv_usec - t0.tv_usec));
  return 0;
}

void dump_coo(const std::string &tag, const SMatrix &A) {
  std::vector<int32_t> row, col;
  std::vector<float> val;
  for (int k = 0; k < A.outerSize(); ++k)
    for (SMatrix::InnerIterator it(A, k); it; ++it) {
      row.push_back((int32_t)it.row());
      col.push_back((int32_t)it.col());
      val.push_back(it.value());
    }
  npy_i32(tag + "_row", row);
  npy_i32(tag + "_col", col);
  npy_f32(tag + "_val", val);
}

std::vector<float> vecf(const VectorXf &v) { return std::vector<float>(v.data(), v.data() + v.size()); }

int mode_qp(int argc, char **argv) {
  BookshelfParser parser;
  PlaceDB *db = new PlaceDB();
  load_design(argc, argv, parser, db);
  if (arg_flag(argc, argv, "timeOnly")) {
    // wall time of quadraticPlacement alone (the reference's own per-
    // iteration lines on stdout carry the per-iteration times)
    QPPlacer qp(db);
    double t0 = now();
    qp.quadraticPlacement();
    scalar("qp_seconds", now() - t0);
    scalar("qp_hpwl", db->calcHPWL());
    return 0;
  }
  dump_design(db);
  if (arg_flag(argc, argv, "center")) db->moveNodesCenterToCenter();
  std::string cin = arg_str(argc, argv, "centers", "");
  if (!cin.empty()) {
    std::vector<float> xy(db->dbNodes.size() * 2);
    FILE *f = fopen(cin.c_str(), "rb");
    if (!f || fread(xy.data(), 4, xy.size(), f) != xy.size()) {
      fprintf(stderr, "bad centers file\n");
      return 2;
    }
    fclose(f);
    for (size_t i = 0; i < db->dbNodes.size(); ++i) db->setModuleCenter_2D(db->dbNodes[i], xy[2 * i], xy[2 * i + 1]);
  }
  npy_f32("qp_matrix_positions", positions(db->dbNodes), 2);
  scalar("qp_matrix_hpwl", db->calcNetBoundPins());
  QPPlacer qp(db);
  int n = (int)db->dbNodes.size();
  SMatrix XA(n, n), YA(n, n);
  VectorXf Xx(n), Yx(n), Xb(n), Yb(n);
  qp.createSparseMatrix(XA, YA, Xx, Yx, Xb, Yb);
  dump_coo("qp_ax", XA);
  dump_coo("qp_ay", YA);
  npy_f32("qp_bx", vecf(Xb));
  npy_f32("qp_by", vecf(Yb));
  npy_f32("qp_guess_x", vecf(Xx));
  npy_f32("qp_guess_y", vecf(Yx));
  double t0 = now();
  qp.quadraticPlacement();
  scalar("qp_seconds", now() - t0);
  npy_f32("qp_positions", positions(db->dbNodes), 2);
  scalar("qp_hpwl", db->calcHPWL());
  return 0;
}

int mode_mlg(int argc, char **argv) {
  BookshelfParser parser;
  PlaceDB *db = new PlaceDB();
  load_design(argc, argv, parser, db);
  dump_design(db);
  std::string cin = arg_str(argc, argv, "centers", "");
  if (!cin.empty()) {
    std::vector<float> xy(db->dbNodes.size() * 2);
    FILE *f = fopen(cin.c_str(), "rb");
    if (!f || fread(xy.data(), 4, xy.size(), f) != xy.size()) {
      fprintf(stderr, "bad centers file\n");
      return 2;
    }
    fclose(f);
    for (size_t i = 0; i < db->dbNodes.size(); ++i) db->setModuleCenter_2D(db->dbNodes[i], xy[2 * i], xy[2 * i + 1]);
  }
  long long skip = atoll(arg_str(argc, argv, "randSkip", "0").c_str());
  for (long long k = 0; k < skip; ++k) rand();
  float td = (float)atof(arg_str(argc, argv, "targetDensity", "1.0").c_str());
  npy_f32("input_positions", positions(db->dbNodes), 2);
  scalar("input_hpwl", db->calcHPWL());
  SAMacroLegalizer legal(db);
  legal.setTargetDensity(td);
  double t0 = now();
  legal.legalization();
  scalar("mlg_seconds", now() - t0);
  npy_f32("mlg_positions", positions(db->dbNodes), 2);
  scalar("mlg_hpwl", db->calcHPWL());
  scalar("mlg_overlap", legal.totalMacroOverlap);
  scalar("mlg_overlap_free", legal.overlapFree ? 1 : 0);
  scalar("mlg_total_hpwl", legal.totalHPWL);
  scalar("mlg_cell_covered", legal.totalCellAreaCovered);
  scalar("mlg_bins", legal.binDimension.x);
  scalar("mlg_macros", (double)legal.dbMacros.size());
  scalar("mlg_rand_next", rand());
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: ref_harness gp|fft|time ...\n");
    return 2;
  }
  gArg.Init(argc, argv);
  g_out = arg_str(argc, argv, "out", ".");
  mkdir(g_out.c_str(), 0755);
  // The reference logs to stdout; keep the harness's stdout for that and
  // write results to files only.
  g_scalars.open(g_out + "/scalars.txt");
  std::string mode = argv[1];
  if (mode == "gp") return mode_gp(argc, argv);
  if (mode == "fft") return mode_fft(argc, argv);
  if (mode == "time") return mode_time(argc, argv);
  if (mode == "legal") return mode_legal(argc, argv);
  if (mode == "qp") return mode_qp(argc, argv);
  if (mode == "mlg") return mode_mlg(argc, argv);
  if (mode == "dp") return mode_dp(argc, argv);
  fprintf(stderr, "unknown mode %s\n", mode.c_str());
  return 2;
}
