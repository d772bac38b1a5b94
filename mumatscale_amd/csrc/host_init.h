// Host-only part of the CA step library: the reference's RNG stream and
// nucleation set-up, and the folding of step constants.  Built with g++ so
// that libm calls and float arithmetic are those of the gcc-built reference.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "mmca_internal.h"

namespace mmca {

// glibc rand(): TYPE_3 additive feedback generator x[i] = x[i-3] + x[i-31]
// (published algorithm of glibc srandom_r/random_r), seeded as the reference
// does with srand(iproc + 1 + seed) (ca_main.c:215).
struct GlibcRand {
    int32_t r[31];
    int f = 3, b = 0;
    bool gauss_have = false;   // functions.c:380 static have_one_stored
    float gauss_next = 0.0f;   // functions.c:381 static next_val
    void seed(unsigned int s);
    int32_t next();
    float scale();             // getRandScale (functions.c:441-445)
    float gaussian();          // getRandomGaussian (functions.c:377-405)
};

// nucleation_t (globals.h:147-160)
struct NucPt {
    double x, y, z;
    size_t cx, cy, cz;
    double gx, gy, gz;
    double threshold;
};

// Nucleation sites of one subblock in raster (padded index) order.
struct SitePlan {
    std::vector<NucPt> nuc_pt;       // FixedNuc + NucVolume points of the subblock
    std::vector<int> site_idx;       // padded cell index
    std::vector<float> site_thr;     // undercooling threshold
    std::vector<int> site_fixed;     // first matching nuc_pt entry, -1 if none
    std::vector<int> site_ls;        // AM lsindex at set-up
    std::vector<double> fixed_rot;   // 9 per nuc_pt entry
};

long long padded_cells(const mmca_params &p);
int build_consts(const mmca_params &p, Consts &c);
double host_temp(const Consts &c, double sim_time, int x, int y, int z);
void rotmat_of(double a0, double a1, double a2, double *R);
int plan_nucleation(const mmca_params &params, const Consts &c, const mmca_nucleation_cfg *cfg,
                    uint64_t timestep, GlibcRand &rng, SitePlan &out);
// next n random orientations (setNucOrientation, grain.c:118-129)
void draw_orientations(GlibcRand &rng, int oriented, int n, double *angles3, double *rot9);

}  // namespace mmca
