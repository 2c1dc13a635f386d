// ensemble.hip -- many independent boxes advanced together on one GPU.
//
// Each box is a full context with its own stream and its own captured
// timestep graph.  A step of the ensemble issues every box's step before
// waiting for any, so the boxes' kernels overlap on the GPU (small boxes
// cannot fill 256 CUs alone).  Boxes share nothing; a box stopped by the
// CFL check stays frozen while the others continue.
#include "internal.hpp"

#include <string>
#include <vector>

struct mhit_ensemble {
    std::vector<mhit_ctx*> boxes;
};

namespace mhit {
int step_async(mhit_ctx* c, int nsteps);     // api.hip: enqueue without waiting
int finish_steps(mhit_ctx* c);               // api.hip: wait, read the scalars back
}

using namespace mhit;

extern "C" {

int mhit_ensemble_create(int32_t nbox, const mhit_params* params, mhit_ensemble** out) {
    if (!out || nbox <= 0 || !params) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    *out = nullptr;
    mhit_ensemble* e = new mhit_ensemble();
    for (int32_t b = 0; b < nbox; b++) {
        mhit_ctx* c = nullptr;
        int rc = mhit_create(&params[b], &c);
        if (rc) {
            set_error("box " + std::to_string(b) + ": " + mhit_last_error());
            mhit_ensemble_destroy(e);
            return rc;
        }
        e->boxes.push_back(c);
    }
    *out = e;
    return MHIT_OK;
}

int mhit_ensemble_box(mhit_ensemble* e, int32_t i, mhit_ctx** box) {
    if (!e || !box || i < 0 || i >= (int32_t)e->boxes.size()) { set_error("bad box index"); return MHIT_ERR_ARG; }
    *box = e->boxes[i];
    return MHIT_OK;
}

int mhit_ensemble_step(mhit_ensemble* e, int nsteps) {
    if (!e || nsteps < 0) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    int rc;
    // interleave the boxes step by step so that no stream runs far ahead
    // (a box already stopped on the CFL check is skipped, not an error)
    for (int s = 0; s < nsteps; s++)
        for (mhit_ctx* c : e->boxes)
            if (!c->S_host.halted && (rc = step_async(c, 1))) return rc;
    int status = MHIT_OK;
    for (mhit_ctx* c : e->boxes) {
        rc = finish_steps(c);
        if (rc == MHIT_ERR_CFL) status = MHIT_ERR_CFL;
        else if (rc) return rc;
    }
    if (status == MHIT_ERR_CFL) set_error("at least one box stopped on cou > 7");
    return status;
}

int mhit_ensemble_destroy(mhit_ensemble* e) {
    if (!e) return MHIT_OK;
    for (mhit_ctx* c : e->boxes) mhit_destroy(c);
    delete e;
    return MHIT_OK;
}

}  // extern "C"
