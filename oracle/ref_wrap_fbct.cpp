// extern "C" shim over the reference's own 2D CPU kernels (geom/fbct_os/fbct_proj.c,
// geom/fbct_os/fbct_back.c), compiled from /root/reference by oracle/Makefile into
// oracle/_ref/.  TEST INFRASTRUCTURE ONLY: it lets the tests check the oracle's
// restatement against the reference's compiled code.  It declares nothing of its
// own beyond this dispatcher; the declarations come from the reference headers.
#include <cstring>
#include "fbct_back.h"
#include "fbct_proj.h"

extern "C" int ref_fbct(const char* op, const int* nxy, const float* dxy, const float* oxy, int nu,
                        float du, float ou, float sad, float add, float sdd, int nviews, double* betas,
                        int dettype, float* in, float* out) {
    if (!std::strcmp(op, "proj,rd"))
        return !fbct_proj_rd(nxy[0], nxy[1], dxy[0], dxy[1], oxy[0], oxy[1], nu, du, ou, sad, add, sdd,
                             nviews, betas, dettype, in, out);
    if (!std::strcmp(op, "proj,dd"))
        return !fbct_proj_dd(nxy[0], nxy[1], dxy[0], dxy[1], oxy[0], oxy[1], nu, du, ou, sad, add, sdd,
                             nviews, betas, dettype, in, out);
    if (!std::strcmp(op, "proj,tf"))
        return !fbct_proj_tf(nxy[0], nxy[1], dxy[0], dxy[1], oxy[0], oxy[1], nu, du, ou, sad, add, sdd,
                             nviews, betas, dettype, in, out);
    if (!std::strcmp(op, "back,dd"))
        return !fbct_back_dd(nxy[0], nxy[1], dxy[0], dxy[1], oxy[0], oxy[1], nu, du, ou, sad, add, sdd,
                             nviews, betas, dettype, out, in);
    if (!std::strcmp(op, "back,tf"))
        return !fbct_back_tf(nxy[0], nxy[1], dxy[0], dxy[1], oxy[0], oxy[1], nu, du, ou, sad, add, sdd,
                             nviews, betas, dettype, out, in);
    if (!std::strcmp(op, "back,pd"))
        return !fbct_back_pd(nxy[0], nxy[1], dxy[0], dxy[1], oxy[0], oxy[1], nu, du, ou, sad, add, sdd,
                             nviews, betas, dettype, out, in);
    return 1;
}
