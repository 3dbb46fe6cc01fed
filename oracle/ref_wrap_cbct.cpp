// extern "C" shim over the reference's own 3D CPU kernels (geom/cbct_os/cbct_proj.c,
// geom/cbct_os/cbct_back.c), compiled from /root/reference by oracle/Makefile into
// oracle/_ref/.  TEST INFRASTRUCTURE ONLY (see ref_wrap_fbct.cpp).
// map == NULL selects the 15-argument no-map family, as cbct_geom_mex does for nrhs == 15.
#include <cstring>
#include "cbct_back.h"
#include "cbct_proj.h"

extern "C" int ref_cbct(const char* op, const int* n, const float* d, const float* o, const int* nuv,
                        const float* duv, const float* ouv, float sad, float add, float sdd, int nviews,
                        double* betas, float* couchz, int dettype, float* in, unsigned char* map,
                        float* out) {
    bool* m = reinterpret_cast<bool*>(map);
#define GEOM n[0], n[1], n[2], d[0], d[1], d[2], o[0], o[1], o[2], nuv[0], nuv[1], duv[0], duv[1], ouv[0], \
             ouv[1], sad, add, sdd, nviews, betas, couchz, dettype
    bool ok = false;
    if (!std::strcmp(op, "proj,rd")) ok = m ? false : cbct_proj_rd(GEOM, in, out);  // no map form (cbct_geom_mex.c:169)
    else if (!std::strcmp(op, "proj,dd")) ok = m ? cbct_proj1_dd(GEOM, in, out, m) : cbct_proj_dd(GEOM, in, out);
    else if (!std::strcmp(op, "proj,tf")) ok = m ? cbct_proj1_tf(GEOM, in, out, m) : cbct_proj_tf(GEOM, in, out);
    else if (!std::strcmp(op, "back,dd")) ok = m ? cbct_back1_dd(GEOM, out, in, m) : cbct_back_dd(GEOM, out, in);
    else if (!std::strcmp(op, "back,tf")) ok = m ? cbct_back1_tf(GEOM, out, in, m) : cbct_back_tf(GEOM, out, in);
    else if (!std::strcmp(op, "back,pd"))
        ok = m ? cbct_back1_pd(GEOM, out, in, m, true) : cbct_back_pd(GEOM, out, in, true);
    else if (!std::strcmp(op, "back,hpd"))
        ok = m ? cbct_back1_pd(GEOM, out, in, m, false) : cbct_back_pd(GEOM, out, in, false);
#undef GEOM
    return ok ? 0 : 1;
}
