// Adapter that runs the reference placer's QP and global-placement stages on
// libeplace_amd from the reference's own PlaceDB.  It belongs next to
// EPlace/eplace.h in the reference tree and compiles against its headers
// (placedb.h and what it includes) plus include/eplace_amd.h.
//
// Replaces, in main/ePlace_main.cpp:
//   lines 121-126  QPPlacer(placedb)->quadraticPlacement()  -> quadraticPlacementAMD
//   lines 141-168  EPlacer_2D + EplaceNesterovOpt + opt()     -> EPlacerAMD + opt()
//   lines 174-186  SAMacroLegalizer(placedb)->legalization()  -> EPlacerAMD::macroLegalization
//   lines 188-219  switch2FillerOnly / opt / switch2cGP / opt -> the same names on EPlacerAMD
//   lines 271-283  DetailedPlacer(placedb)->detailedPlacement() -> detailedPlacementAMD
//
// Numbering: nodes are 0..N-1 in dbNodes order, terminals N..N+T-1 in
// dbTerminals order (the parser gives each list its own idx, parser.cpp:412-433).
// Pin offsets are from the module centre, as Pin::calculateAbsolutePos uses them.
// Every library error becomes a std::runtime_error carrying ep_last_error();
// the library never exits the process.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "eplace_amd.h"
#include "placedb.h"

namespace eplace_amd {

// ep_design over a PlaceDB.  The arrays are copies: the PlaceDB may change
// afterwards (write-back) without invalidating them.
struct DesignArrays {
    explicit DesignArrays(PlaceDB *db) {
        const int32_t N = (int32_t)db->dbNodes.size(), T = (int32_t)db->dbTerminals.size();
        for (Module *m : db->dbNodes) {
            w_.push_back(m->width); h_.push_back(m->height);
            x_.push_back(m->getCenter().x); y_.push_back(m->getCenter().y);
        }
        for (Module *t : db->dbTerminals) {
            tw_.push_back(t->width); th_.push_back(t->height);
            tx_.push_back(t->getCenter().x); ty_.push_back(t->getCenter().y);
            ni_.push_back(t->isNI ? 1 : 0);
        }
        start_.push_back(0);
        for (Net *n : db->dbNets) {
            for (Pin *p : n->netPins) {
                obj_.push_back(objectIndex(db, p->module, N, T));
                dx_.push_back(p->offset.x); dy_.push_back(p->offset.y);
            }
            start_.push_back((int64_t)obj_.size());
        }
        for (SiteRow &r : db->dbSiteRows) {
            rb_.push_back(r.bottom); rh_.push_back(r.height);
            r0_.push_back(r.start.x); r1_.push_back(r.end.x);
        }
        d.num_nodes = N; d.num_terminals = T; d.num_nets = (int32_t)db->dbNets.size();
        d.num_rows = (int32_t)rb_.size(); d.num_pins = (int64_t)obj_.size();
        d.node_w = w_.data(); d.node_h = h_.data(); d.node_x = x_.data(); d.node_y = y_.data();
        d.term_w = tw_.data(); d.term_h = th_.data(); d.term_x = tx_.data(); d.term_y = ty_.data();
        d.term_ni = ni_.data(); d.net_start = start_.data(); d.pin_obj = obj_.data();
        d.pin_dx = dx_.data(); d.pin_dy = dy_.data();
        d.row_bottom = rb_.data(); d.row_height = rh_.data(); d.row_x0 = r0_.data(); d.row_x1 = r1_.data();
        d.common_row_height = db->commonRowHeight;
    }
    DesignArrays(const DesignArrays &) = delete;
    DesignArrays &operator=(const DesignArrays &) = delete;

    ep_design d{};

private:
    static int32_t objectIndex(PlaceDB *db, Module *m, int32_t N, int32_t T) {
        const int32_t k = m->idx;
        if (k >= 0 && k < T && db->dbTerminals[k] == m) return N + k;
        if (k >= 0 && k < N && db->dbNodes[k] == m) return k;
        throw std::runtime_error("eplace_amd: pin on a module outside dbNodes/dbTerminals: " + m->name);
    }
    std::vector<float> w_, h_, x_, y_, tw_, th_, tx_, ty_, dx_, dy_;
    std::vector<uint8_t> ni_;
    std::vector<int64_t> start_;
    std::vector<int32_t> obj_;
    std::vector<double> rb_, rh_, r0_, r1_;
};

// Node centres (dbNodes order, x/y pairs) into the PlaceDB through
// setModuleCenter_2D, which also refreshes the pins' absolute positions.
inline void writeNodeCenters(PlaceDB *db, const std::vector<float> &xy) {
    for (size_t k = 0; k < db->dbNodes.size(); ++k)
        db->setModuleCenter_2D(db->dbNodes[k], xy[2 * k], xy[2 * k + 1]);
}

// QPPlacer(db).quadraticPlacement() (qplace.cpp:20-103) on the GPU: the
// bound-to-bound system is built and solved on the device from the core
// centre, and the database's nodes move to the solution.
inline ep_qp_stats quadraticPlacementAMD(PlaceDB *db, int maxIterations = 20, int device = 0) {
    DesignArrays a(db);
    ep_qp_params q{};
    q.max_iterations = maxIterations;              // -IPiteCount (default 20)
    q.solver_iterations = 100;                     // BiCGSTAB setMaxIterations(100)
    q.target_error = 0.000001f;                    // target_error
    q.solver_tolerance = 1.1920928955078125e-07f;  // Eigen's float epsilon
    q.start_at_center = 1;
    q.device = device;
    std::vector<float> xy(2 * (size_t)a.d.num_nodes);
    std::vector<double> log(4 * (size_t)(maxIterations + 1));
    ep_qp_stats st{};
    if (ep_qplace(&a.d, &q, xy.data(), log.data(), &st) != EP_OK)
        throw std::runtime_error(ep_qp_last_error());
    writeNodeCenters(db, xy);
    return st;
}

// DetailedPlacer(db)->detailedPlacement() (DetailedPlacement/detailed.cpp:74-85)
// in the library on the database's legal node positions (after the
// reference's AbacusLegalizer, as -internalLegal -internalDP run them); the
// nodes move to the result.
inline ep_dp_stats detailedPlacementAMD(PlaceDB *db, int passes = 2) {
    DesignArrays a(db);
    ep_dp_params p{};
    p.site_step = db->dbSiteRows.front().step;
    p.passes = passes;
    std::vector<float> xy(2 * (size_t)a.d.num_nodes);
    ep_dp_stats st{};
    if (ep_detailed_place(&a.d, nullptr, &p, xy.data(), &st) != EP_OK)
        throw std::runtime_error(ep_host_last_error());
    writeNodeCenters(db, xy);
    return st;
}

// Step rules of EplaceNesterovOpt: -bktrk, -bb, or the learned controller.
enum class Step { Vanilla = 0, Backtracking = 1, Learned = 2, BarzilaiBorwein = 3 };

// EPlacer_2D + EplaceNesterovOpt on the GPU.  The placer keeps the fillers,
// the stage (mGP / FILLERONLY / cGP), the penalty factor and its own rand()
// stream between calls, as the reference's objects do between the blocks of
// main/ePlace_main.cpp.
class EPlacerAMD {
public:
    // EPlacer_2D(db) + setTargetDensity + initialization() (eplace.cpp:39-50)
    EPlacerAMD(PlaceDB *db, float targetDensity, bool lse = false, int device = 0)
        : db_(db), a_(db), device_(device) {
        ep_params p{};
        p.target_density = targetDensity;
        p.wirelength_model = lse ? 1 : 0;   // -LSE
        p.filler_seed = 1;                  // an unseeded process's rand()
        p.filler_count = -1;                // the reference's filler count
        p.poisson_mode = 2;                 // split-FP16 transforms
        p.device = device;
        check(ep_create(&a_.d, &p, &ctx_));
        check(ep_total_gradient_update(ctx_));
        check(ep_penalty_factor_init(ctx_));
    }
    ~EPlacerAMD() {
        if (ctrl_) ep_ctrl_destroy(ctrl_);
        ep_destroy(ctx_);
    }
    EPlacerAMD(const EPlacerAMD &) = delete;
    EPlacerAMD &operator=(const EPlacerAMD &) = delete;

    // Step::Learned needs the controller's weights (EP_CTRL_PARAMS floats,
    // eplace_amd/controller_params.npy in the library's tree).
    void setController(const float *params) {
        if (ctrl_) check(ep_ctrl_set_params(ctrl_, params), ep_ctrl_last_error(ctrl_));
        else if (ep_ctrl_create(device_, params, 1, &ctrl_) != EP_OK)
            throw std::runtime_error(ctrl_ ? ep_ctrl_last_error(ctrl_) : "eplace_amd: ep_ctrl_create failed");
    }

    // opt->opt() for the current stage (opt.hpp:29-36), then the node centres
    // go back into the PlaceDB.  Stops at targetOverflow (mGP 0.10) or after
    // MAX_ITERATION = 1000 iterations.
    ep_run_stats opt(Step step = Step::Vanilla, float targetOverflow = 0.10f) {
        if (step == Step::Learned && !ctrl_)
            throw std::runtime_error("eplace_amd: Step::Learned without setController()");
        ep_opt_params o{};
        o.method = (int32_t)step;
        o.target_overflow = targetOverflow;
        o.max_iterations = 1000;
        o.max_backtracks = 30;
        o.sync_every = 16;
        o.controller = step == Step::Learned ? ctrl_ : nullptr;
        const bool mgp = stats().stage == EP_STAGE_MGP;
        ep_run_stats s{};
        check(ep_nesterov_run(ctx_, &o, &s));
        if (mgp) mGPIterationCount = s.iterations;
        writeBack();
        return s;
    }
    // EPlacer_2D::switch2FillerOnly / switch2cGP (eplace.cpp:887-905); cGP
    // starts from lambda / 1.1^(0.1 * mGPIterationCount), the count of the
    // last mGP run.
    void switch2FillerOnly() { check(ep_switch_to_filler_only(ctx_)); }
    void switch2cGP() { check(ep_switch_to_cgp(ctx_, mGPIterationCount)); }
    // SAMacroLegalizer(db) + setTargetDensity + legalization()
    // (legalizer.cpp:701-790) on the placer's positions and rand() stream.
    ep_mlg_stats macroLegalization(float targetDensity) {
        ep_mlg_params m{};
        m.target_density = targetDensity;
        m.rand_seed = 1;
        ep_mlg_stats st{};
        check(ep_macro_legalize(ctx_, &a_.d, &m, &st));
        writeBack();
        return st;
    }
    ep_stats stats() {
        ep_stats s{};
        check(ep_get_stats(ctx_, &s));
        return s;
    }
    ep_ctx *context() { return ctx_; }
    int mGPIterationCount = 0;

private:
    void writeBack() {
        ep_info i{};
        check(ep_get_info(ctx_, &i));
        std::vector<float> xy(2 * (size_t)i.num_objects);
        check(ep_read(ctx_, EP_ARR_POSITIONS, xy.data(), xy.size()));
        writeNodeCenters(db_, xy);
    }
    void check(int rc, const char *msg = nullptr) {
        if (rc != EP_OK) throw std::runtime_error(msg ? msg : ep_last_error(ctx_));
    }
    PlaceDB *db_;
    DesignArrays a_;
    int device_;
    ep_ctx *ctx_ = nullptr;
    ep_ctrl *ctrl_ = nullptr;
};

}  // namespace eplace_amd
