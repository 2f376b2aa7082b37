/* gravis_oracle.h -- declarations shared by the CPU oracle's translation
 * units (test infrastructure only). */
#ifndef GRAVIS_ORACLE_H
#define GRAVIS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

typedef struct orc_rng {
    uint32_t key[8];
    uint64_t counter;   /* next block to generate */
    uint32_t buf[64];   /* 4 blocks, as rand_chacha buffers them */
    int idx;            /* next u32 word in buf; 64 = empty */
} orc_rng;

void orc_rng_seed_u64(orc_rng* r, uint64_t state);
uint64_t orc_rng_next_u64(orc_rng* r);
double orc_rng_f64(orc_rng* r);
double orc_rng_range(orc_rng* r, double lo, double hi);
void orc_root_box(size_t n, const double* x, const double* y, const double* z, double out[4]);
void orc_bf_accel(size_t n, double eps, const double* x, const double* y, const double* z,
                  const double* m, double* ax, double* ay, double* az);
void orc_shift_to_com(size_t n, double* x, double* y, double* z, double* vx, double* vy,
                      double* vz, const double* m);
void orc_bh_accel(size_t n, double eps, double theta, const double* x, const double* y,
                  const double* z, const double* m, double* ax, double* ay, double* az,
                  int nthreads);
uint64_t orc_bh_group_accel(size_t n, double eps, double theta, uint32_t leaf, uint32_t group,
                            uint32_t fcap, int split, const double* x, const double* y,
                            const double* z, const double* m, double* ax, double* ay, double* az,
                            int nthreads);
uint64_t orc_bh_group_accel_eps(size_t n, const double* eps_i, double theta, uint32_t leaf,
                                uint32_t group, uint32_t fcap, const double* x, const double* y,
                                const double* z, const double* m, double* ax, double* ay,
                                double* az, int nthreads);
#endif
