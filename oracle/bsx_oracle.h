/*
 * bsx_oracle.h — CPU restatement of Bsoft's Fourier-space reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline — never as the product path.
 *
 * Built twice from bsx_oracle.c: liboracle64.so (FP64 data, the parity oracle)
 * and liboracle32.so (FP32 images/accumulators with double intermediates and
 * OpenMP over partial maps, like breconstruct -threads T: the CPU baseline).
 */
#ifndef BSX_ORACLE_H
#define BSX_ORACLE_H

#include "../include/bsx.h"

#ifdef OR_FP32
typedef float orr;
#else
typedef double orr;
#endif

#ifdef __cplusplus
extern "C" {
#endif

int    or_real_bytes(void);
int    or_part_ft_size(int xsize, double scale, int pad_factor);
int    or_find_next_power(int start, int power);
void   or_view_matrix(const double view[4], double mat[9]);
double or_lambda(double volt);
double or_check_resolution(double res, int n, double sampling, int ndim);
double or_ctf_value(const bsx_particle* p, double sx, double sy, int flip, double wiener);
int    or_fft(orr* data, long nx, long ny, long nz, long batch, int dir, int norm);
double or_background(const float* img, long box, const double hdr_ori[2]);
void   or_rescale_f32(float* img, long n, double nuavg, double nustd);
int    or_rescale_to_avg_std(orr* data, long n, double nuavg, double nustd);
int    or_filter_sinc(orr* data, long nx, long ny, long nz, const double origin[3]);
int    or_stage_particle(const bsx_recon_params* rp, const bsx_particle* p, const float* img,
                         orr* out_ft, double* part_weight);

typedef struct or_recon or_recon;
or_recon* or_recon_new(const bsx_recon_params* rp, int npartial);
void   or_recon_free(or_recon* r);
int    or_recon_insert(or_recon* r, const bsx_particle* parts, const float* imgs, long n,
                       int nthreads);
int    or_recon_partial_acc(or_recon* r, int ipartial, orr* planes);
int    or_recon_set_partial_acc(or_recon* r, int ipartial, const orr* planes, long count);
int    or_recon_finish(or_recon* r, int imap, int nmaps, orr* map_out, orr* fom_out,
                       orr* ft_out, bsx_stats* st);

int    or_fsc(const orr* a, const orr* b, long n, double sampling, double hi_res,
              double* fsc, int cap);
int    or_fsi_kernel(int width, int power, double* table);
int    or_map_transform(const orr* map, long n, const double origin[3], orr* ft_out);
int    or_central_section(const orr* ft, long n, double sampling, const double mat[9],
                          double resolution, int kw, int kp, int center, orr* out);
double or_recip_space_fom(const orr* part, const orr* ref, long n, double sampling,
                          double hi_res, double lo_res, double* cv_out);
long   or_asu_views_c1(double theta_step, double phi_step, int flag, double* views, long cap);
double or_friedel(const orr* ft, long nx, long ny, long nz);

#ifdef __cplusplus
}
#endif

#endif
