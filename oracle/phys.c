/* Oracle: FP32 and FP64 instantiations of the physical estimator
 * (phys_impl.h), the FP64 path-replay adjoint and the camera basis.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R float
#define SFX(x) x##_f32
#include "phys_impl.h"
#undef R
#undef SFX
#undef VT
#undef ST

#define R double
#define SFX(x) x##_f64
#include "phys_impl.h"
#undef R
#undef SFX
#undef VT
#undef ST

/* Camera basis, computed in double and rounded once to float:
 *   w  = normalize(target - eye)              (forward)
 *   du = normalize(cross(w, up)) * tan(vfov/2) * W/H
 *   dv = cross(u, w) * tan(vfov/2)
 * A pixel (x + jx, y + jy) maps to normalize(du*fx + dv*fy + w) with
 * fx = 2(x+jx)/W - 1, fy = 1 - 2(y+jy)/H (row 0 at the top). */
void or_camera_basis(const float eye[3], const float target[3], const float up[3], float vfov_deg, int W, int H,
                     or_cam *cam) {
	double w[3] = {(double)target[0] - eye[0], (double)target[1] - eye[1], (double)target[2] - eye[2]};
	double lw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
	for (int k = 0; k < 3; k++)
		w[k] /= lw;
	double u[3] = {w[1] * up[2] - w[2] * up[1], w[2] * up[0] - w[0] * up[2], w[0] * up[1] - w[1] * up[0]};
	double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
	for (int k = 0; k < 3; k++)
		u[k] /= lu;
	double v[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
	double th = tan(0.5 * (double)vfov_deg * 3.14159265358979323846 / 180.0);
	double aspect = (double)W / (double)H;
	for (int k = 0; k < 3; k++) {
		cam->eye[k] = eye[k];
		cam->w[k] = (float)w[k];
		cam->du[k] = (float)(u[k] * th * aspect);
		cam->dv[k] = (float)(v[k] * th);
	}
}

/* Adjoint (FP64). For each sample the path is re-traced from its Philox
 * counters; vertex records (triangle, albedo, NEE term X, throughput, 1/q)
 * live only for that one path. With H_n = 0 the reverse sweep
 *     D_i = X_i + H_{i+1} / q_i,  dI/da_i += beta_i * D_i,  H_i = a_i * D_i
 * is the exact derivative of the sample's value with respect to the albedo at
 * vertex i (beta_i depends only on earlier vertices); emission derivatives are
 * the NEE coefficients beta*a*Y and 1 for the emitter seen at depth 0. */
void or_phys_adjoint_f64(const void *sp, const or_cam *cam, const or_params *p, const float *dl_dimage, int nthreads,
                         double *g_albedo, double *g_emission) {
	const Scene_f64 *s = (const Scene_f64 *)sp;
	int npix = p->width * p->height;
	int nt = nthreads > 0 ? nthreads : 1;
	size_t n3 = 3 * (size_t)(s->n > 0 ? s->n : 1);
	double *ga = (double *)calloc((size_t)nt * n3, sizeof(double));
	double *ge = (double *)calloc((size_t)nt * n3, sizeof(double));
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
	for (int i = 0; i < npix; i++) {
		int tid = 0;
#ifdef _OPENMP
		tid = omp_get_thread_num();
#endif
		double *lga = ga + (size_t)tid * n3, *lge = ge + (size_t)tid * n3;
		/* every term below is a multiple of this pixel's weight: a zero
		 * weight contributes exactly nothing, so its paths are not traced */
		if (dl_dimage[3 * i] == 0.0f && dl_dimage[3 * i + 1] == 0.0f && dl_dimage[3 * i + 2] == 0.0f)
			continue;
		int x = i % p->width, y = i / p->width;
		V_f64 wgt = vmk_f64(dl_dimage[3 * i] / p->spp, dl_dimage[3 * i + 1] / p->spp, dl_dimage[3 * i + 2] / p->spp);
		Vtx_f64 rec[64];
		or_counts cnt = {0, 0, 0};
		or_params pp = *p;
		if (pp.max_depth > 64)
			pp.max_depth = 64;
		for (int k = 0; k < p->spp; k++) {
			int nrec = 0;
			path_f64(s, cam, &pp, x, y, (uint32_t)k, &cnt, rec, &nrec, wgt, lge);
			V_f64 H = vmk_f64(0, 0, 0);
			for (int j = nrec - 1; j >= 0; j--) {
				V_f64 D = vadd_f64(rec[j].X, vs_f64(H, rec[j].invq));
				V_f64 g = vmul_f64(vmul_f64(wgt, rec[j].beta), D);
				int t = rec[j].tri;
				lga[3 * t] += g.x;
				lga[3 * t + 1] += g.y;
				lga[3 * t + 2] += g.z;
				H = vmul_f64(rec[j].a, D);
			}
		}
	}
	for (int t = 0; t < nt; t++)
		for (size_t j = 0; j < 3 * (size_t)s->n; j++) {
			g_albedo[j] += ga[(size_t)t * n3 + j];
			g_emission[j] += ge[(size_t)t * n3 + j];
		}
	free(ga);
	free(ge);
}
