/* photic_amd -- C host code: photic_invert_ordered, the reference's pixel ordering, look-up
 * table and hot start (model/samodel.c:649-660, 701-814, 980-1109) recast as batches on the
 * GPU.  The kernels and the method are described in ph_order.hip; here they are driven:
 *
 *   1. every pixel's spectral index, a stable sort of the modelled pixels by it;
 *   2. the leaders (sorted positions 0, run, 2 run, ...) inverted with the full depth list;
 *   3. every follower's role: a LUT match with one of the `window + 1` leaders up to its own,
 *      or a hot start from its own leader's depth;
 *   4. the LUT matches take their leader's parameter vector (one evaluation each);
 *   5. the hot starts inverted with one start each, then, when a restart level is set, those
 *      still above it inverted again with the full depth list (model/samodel.c:1013-1031,
 *      compiled in the reference only with RESTART, which it leaves at 0, :98);
 *   6. everything scattered back to pixel order.
 *
 * Each batch is one photic_invert_h launch, so every pixel's result is the single-date
 * kernel's for its own start rules; the counts between the steps come back to the host. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "photic.h"
#include "ph_launch.h"

void photic_order_cfg_default(photic_order_cfg *o)
{
    memset(o, 0, sizeof(*o));
    o->run = 64;
    o->window = 16;
    o->use_lut = 1;
    o->match_max = 1.5;      /* lut_match_threshold for one scene, model/samodel.c:896 */
    o->entry_max = 1.5;      /* lut_entry_threshold, :897 */
    o->restart_misfit = 0.0; /* no restart: the reference compiles RESTART out (:98) */
}

static int od_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return PHOTIC_OK;
    (void)hipGetLastError();
    char msg[256];
    snprintf(msg, sizeof(msg), "%s: %s", what, hipGetErrorString(e));
    return ph_error(PHOTIC_EDEVICE, msg);
}

static int od_launch(int rc, const char *what)
{
    if (rc == 0) return PHOTIC_OK;
    if (rc == PH_LAUNCH_UNSUPPORTED) return ph_error(PHOTIC_EUNSUPPORTED, what);
    return od_hip((hipError_t)rc, what);
}

/* Device outputs of m pixels (every field of photic_outputs) in one allocation. */
typedef struct {
    void *mem;
    photic_outputs o;
} od_outs;

static int od_outs_alloc(od_outs *b, long m, int n_params)
{
    memset(b, 0, sizeof(*b));
    if (m <= 0) return PHOTIC_OK;
    const size_t f = sizeof(float) * (size_t)m, d = sizeof(double) * (size_t)m;
    const size_t i = sizeof(int) * (size_t)m;
    const size_t bytes = 10 * f + (size_t)(n_params + 4) * d + 4 * i;
    int rc = od_hip(hipMalloc(&b->mem, bytes), "hipMalloc");
    if (rc) return rc;
    char *p = (char *)b->mem;
    b->o.x = (double *)p;
    p += (size_t)n_params * d;
    b->o.rrs_model = (double *)p;
    p += 4 * d;
    b->o.depth = (float *)p;
    b->o.rrs_error = (float *)(p + f);
    b->o.model_error = (float *)(p + 2 * f);
    b->o.albedo = (float *)(p + 3 * f);
    b->o.k_min = (float *)(p + 4 * f);
    b->o.iod = (float *)(p + 5 * f);
    b->o.bottom_type = (float *)(p + 6 * f);
    b->o.b_pct = (float *)(p + 7 * f);
    p += 10 * f;
    b->o.status = (int *)p;
    b->o.n_evals = (int *)(p + i);
    b->o.n_iter = (int *)(p + 2 * i);
    b->o.n_starts = (int *)(p + 3 * i);
    return PHOTIC_OK;
}

static void od_outs_free(od_outs *b)
{
    if (b->mem) hipFree(b->mem);
    b->mem = NULL;
}

static int od_count(const unsigned int *d_count, unsigned int *out, hipStream_t st)
{
    int rc = od_hip(hipMemcpyAsync(out, d_count, sizeof(unsigned int), hipMemcpyDeviceToHost, st),
                    "count read");
    if (!rc) rc = od_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    return rc;
}

int photic_invert_ordered(const photic_cfg *cfg, const photic_order_cfg *ocfg, long n_pix,
                          const float *rrs, const float *sec_sun, const photic_outputs *out,
                          int *role, photic_order_stats *stats, void *stream)
{
    if (!cfg || !ocfg || !out) return ph_error(PHOTIC_EINVAL, "cfg, ocfg and out are required");
    if (n_pix < 0 || n_pix > 2147483647L)
        return ph_error(PHOTIC_EINVAL, "n_pix must be in 0 .. 2^31 - 1");
    if (ocfg->run < 1 || ocfg->window < 0)
        return ph_error(PHOTIC_EINVAL, "run >= 1 and window >= 0 required");
    if (cfg->n_bands != 4 || cfg->precision != PHOTIC_FP64)
        return ph_error(PHOTIC_EUNSUPPORTED, "the ordered inversion is compiled for 4 bands, FP64");
    if (n_pix > 0 && !rrs) return ph_error(PHOTIC_EINVAL, "rrs is NULL");
    if (stats) memset(stats, 0, sizeof(*stats));
    PhInvertArgs a;
    memset(&a, 0, sizeof(a));
    int rc = ph_build_const(cfg, 1, 0, &a.c);
    if (rc) return rc;
    if (n_pix == 0) return PHOTIC_OK;
    const int nbot = a.c.mode ? 0 : a.c.n_bottoms;
    const int n_params = photic_n_params(cfg);
    hipStream_t st = (hipStream_t)stream;
    const double match_max = ocfg->match_max > 0.0 ? ocfg->match_max : 1.5;
    const double entry_max = ocfg->entry_max > 0.0 ? ocfg->entry_max : match_max;
    const double restart = ocfg->restart_misfit; /* <= 0: no restart */
    int noise = 1;
    for (int b = 0; b < 4; b++)
        if (!(ocfg->r_sigma[b] >= 1.0e-6)) noise = 0; /* model/samodel.c:590-598 */

    /* the pixel outputs and roles */
    a.n_pix = n_pix;
    a.rrs = rrs;
    a.sec_sun = sec_sun;
    a.depth = out->depth;
    a.rrs_error = out->rrs_error;
    a.model_error = out->model_error;
    a.albedo = out->albedo;
    a.k_min = out->k_min;
    a.iod = out->iod;
    a.b_pct = out->b_pct;
    a.bottom_type = out->bottom_type;
    a.x_out = out->x;
    a.status = out->status;
    a.n_evals = out->n_evals;
    a.n_iter = out->n_iter;
    a.n_starts = out->n_starts;
    a.rrs_model = out->rrs_model;

    float *key = NULL, *key_s = NULL, *l_rrs_in = NULL, *l_ss = NULL, *g_rrs = NULL, *g_ss = NULL;
    float *h_s = NULL, *g_h = NULL;
    int *pix = NULL, *order = NULL, *role_s = NULL, *src_s = NULL, *list = NULL, *list2 = NULL;
    int *flag = NULL;
    unsigned int *cnt = NULL;
    void *tmp = NULL;
    size_t tmp_bytes = 0;
    od_outs L, H, R;
    memset(&L, 0, sizeof(L));
    memset(&H, 0, sizeof(H));
    memset(&R, 0, sizeof(R));
    unsigned int n_valid = 0, m1 = 0, m2 = 0, m3 = 0;
    const size_t n = (size_t)n_pix;

    rc = od_hip(hipMalloc((void **)&key, sizeof(float) * n), "hipMalloc");
    if (!rc) rc = od_hip(hipMalloc((void **)&key_s, sizeof(float) * n), "hipMalloc");
    if (!rc) rc = od_hip(hipMalloc((void **)&pix, sizeof(int) * n), "hipMalloc");
    if (!rc) rc = od_hip(hipMalloc((void **)&order, sizeof(int) * n), "hipMalloc");
    if (!rc) rc = od_hip(hipMalloc((void **)&cnt, sizeof(unsigned int) * 4), "hipMalloc");
    if (!rc) rc = od_launch(ph_launch_zero_u64((unsigned long long *)cnt, 2, st), "zero counts");
    /* 1. spectral index and the stable sort */
    if (!rc)
        rc = od_launch(ph_launch_order_key(&a, noise, ocfg->r_sigma, key, pix, cnt, role, st),
                       "spectral index kernel");
    if (!rc) rc = od_launch(ph_order_sort(key, key_s, pix, order, n_pix, NULL, &tmp_bytes, st),
                            "sort size");
    if (!rc) rc = od_hip(hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1), "hipMalloc");
    if (!rc) rc = od_launch(ph_order_sort(key, key_s, pix, order, n_pix, tmp, &tmp_bytes, st),
                            "sort");
    if (!rc) rc = od_count(cnt, &n_valid, st);
    const long nv = (long)n_valid;
    const long run = ocfg->run;
    const long n_lead = (nv + run - 1) / run;

    /* 2. the leaders, with the full depth list */
    if (!rc && nv > 0) {
        rc = od_hip(hipMalloc((void **)&l_rrs_in, sizeof(float) * 4 * (size_t)n_lead), "hipMalloc");
        if (!rc && sec_sun)
            rc = od_hip(hipMalloc((void **)&l_ss, sizeof(float) * (size_t)n_lead), "hipMalloc");
        if (!rc) rc = od_outs_alloc(&L, n_lead, n_params);
        if (!rc)
            rc = od_launch(ph_launch_order_gather(n_lead, NULL, run, order, rrs, sec_sun, n_pix,
                                                  NULL, l_rrs_in, l_ss, NULL, st),
                           "leader gather");
        if (!rc)
            rc = photic_invert_h(cfg, n_lead, l_rrs_in, l_ss, NULL, NULL, NULL, &L.o, stream);
        if (!rc)
            rc = od_launch(ph_launch_order_scatter(out, &L.o, n_lead, NULL, run, order, n_pix,
                                                   n_params, 0, st),
                           "leader scatter");
    }
    /* 3. the followers' roles */
    if (!rc && nv > 0) {
        rc = od_hip(hipMalloc((void **)&role_s, sizeof(int) * (size_t)nv), "hipMalloc");
        if (!rc) rc = od_hip(hipMalloc((void **)&src_s, sizeof(int) * (size_t)nv), "hipMalloc");
        if (!rc) rc = od_hip(hipMalloc((void **)&h_s, sizeof(float) * (size_t)nv), "hipMalloc");
        if (!rc) rc = od_hip(hipMalloc((void **)&list, sizeof(int) * (size_t)nv), "hipMalloc");
        if (!rc)
            rc = od_launch(ph_launch_order_decide(nv, (int)run, ocfg->window, ocfg->use_lut,
                                                  match_max, entry_max, rrs, n_pix, order,
                                                  L.o.depth, L.o.model_error, L.o.status,
                                                  L.o.rrs_model, n_lead, role_s, src_s, h_s, st),
                           "role kernel");
        if (!rc && role) rc = od_launch(ph_launch_order_roles(nv, NULL, role_s, order, role, st),
                                        "roles");
    }
    /* 4. LUT matches adopt their leader's vector */
    if (!rc && nv > 0) {
        rc = od_launch(ph_launch_order_list(nv, role_s, 1, NULL, NULL, list, cnt + 1, st),
                       "list");
        if (!rc) rc = od_count(cnt + 1, &m1, st);
        if (!rc && m1 > 0)
            rc = od_launch(ph_launch_order_adopt(&a, nbot, m1, list, order, src_s, L.o.x,
                                                 L.o.status, n_lead, st),
                           "adopt kernel");
    }
    /* 5. hot starts, then restarts with the full list */
    if (!rc && nv > 0) {
        rc = od_launch(ph_launch_order_list(nv, role_s, 2, NULL, NULL, list, cnt + 2, st), "list");
        if (!rc) rc = od_count(cnt + 2, &m2, st);
    }
    if (!rc && m2 > 0) {
        rc = od_hip(hipMalloc((void **)&g_rrs, sizeof(float) * 4 * (size_t)m2), "hipMalloc");
        if (!rc && sec_sun) rc = od_hip(hipMalloc((void **)&g_ss, sizeof(float) * m2), "hipMalloc");
        if (!rc) rc = od_hip(hipMalloc((void **)&g_h, sizeof(float) * m2), "hipMalloc");
        if (!rc) rc = od_outs_alloc(&H, m2, n_params);
        if (!rc)
            rc = od_launch(ph_launch_order_gather(m2, list, 0, order, rrs, sec_sun, n_pix, h_s,
                                                  g_rrs, g_ss, g_h, st),
                           "hot-start gather");
        if (!rc) rc = photic_invert_h(cfg, m2, g_rrs, g_ss, NULL, NULL, g_h, &H.o, stream);
        if (!rc)
            rc = od_launch(ph_launch_order_scatter(out, &H.o, m2, list, 0, order, n_pix,
                                                   n_params, 0, st),
                           "hot-start scatter");
    }
    if (!rc && m2 > 0 && restart > 0.0) {
        /* the restart list, as sorted positions */
        rc = od_hip(hipMalloc((void **)&flag, sizeof(int) * m2), "hipMalloc");
        if (!rc) rc = od_hip(hipMalloc((void **)&list2, sizeof(int) * m2), "hipMalloc");
        if (!rc)
            rc = od_launch(ph_launch_order_restart(m2, H.o.model_error, restart, flag, st),
                           "restart flags");
        if (!rc)
            rc = od_launch(ph_launch_order_list(m2, NULL, -1, flag, list, list2, cnt + 3, st),
                           "list");
        if (!rc) rc = od_count(cnt + 3, &m3, st);
    }
    if (!rc && m3 > 0) {
        rc = od_outs_alloc(&R, m3, n_params);
        if (!rc)
            rc = od_launch(ph_launch_order_gather(m3, list2, 0, order, rrs, sec_sun, n_pix, NULL,
                                                  g_rrs, g_ss, NULL, st),
                           "restart gather");
        if (!rc) rc = photic_invert_h(cfg, m3, g_rrs, g_ss, NULL, NULL, NULL, &R.o, stream);
        if (!rc)
            rc = od_launch(ph_launch_order_scatter(out, &R.o, m3, list2, 0, order, n_pix,
                                                   n_params, 1, st),
                           "restart scatter");
        if (!rc && role)
            rc = od_launch(ph_launch_order_roles(m3, list2, NULL, order, role, st), "roles");
    }
    if (!rc) rc = od_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (stats) {
        stats->n_valid = nv;
        stats->n_leaders = n_lead;
        stats->n_lut = m1;
        stats->n_hot = m2;
        stats->n_restarted = m3;
    }
    od_outs_free(&L);
    od_outs_free(&H);
    od_outs_free(&R);
    hipFree(key);
    hipFree(key_s);
    hipFree(pix);
    hipFree(order);
    hipFree(cnt);
    hipFree(tmp);
    hipFree(l_rrs_in);
    hipFree(l_ss);
    hipFree(g_rrs);
    hipFree(g_ss);
    hipFree(g_h);
    hipFree(h_s);
    hipFree(role_s);
    hipFree(src_s);
    hipFree(list);
    hipFree(list2);
    hipFree(flag);
    return rc;
}
