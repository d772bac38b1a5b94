/*
 * Reader of muMatScale control files (the format of read_ctrl.c:235-812)
 * for the C host driver: fills the step parameters (mmca_params) and the
 * nucleation set-up (mmca_nucleation_cfg) of the whole domain as one block,
 * plus the run-loop settings that loop() reads.
 */
#ifndef MMCA_CTRL_H
#define MMCA_CTRL_H

#include "mmca.h"

#define MMCA_VIS_NONE 0
#define MMCA_VIS_VTK  1
#define MMCA_VIS_XDMF 2

typedef struct mmca_run_cfg {
    char     basefilename[256];   /* BaseFileName                         */
    int      nsub[3];             /* NSubblocks                           */
    int      sbdim[3];            /* NCellsPerSB                          */
    int      pad[3];              /* FaceCtrl per axis                    */
    double   cell_size, courant_limit, dt, finish_time, fs_finish;
    int      screenpfreq, data_write_freq, data_write_start;
    int      vis_format;          /* DataFormat: MMCA_VIS_NONE / _VTK / _XDMF */
    int      randomize_grains;    /* RandomizeGrains                      */
    double   t_init, cool_rate, liquidus_temp, solidus_temp;
    uint32_t random_seed;
    double   bulk_density, bulk_mean, bulk_sigma;
    double   surf_density, surf_mean, surf_sigma;
    int      oriented, do_surface;
    int64_t  max_total_grains;
    double   m_solute[3], c_init, d_sol, d_liq, part_coef, fsgrow;
    int      thermal_ana, am_layers, am_tracks;
    double   gradient, grad_coef, velocity, velo_coef;
    double   laser_power, laser_speed, laser_spot, laser_absorp, laser_hatch;
    double   sub_init_temp, th_conductivity, density, specific_heat, laser_start[3];
    mmca_fixed_nuc  *fixed;       /* FixedNuc entries in file order       */
    int      n_fixed, cap_fixed;
    mmca_nuc_volume *volumes;     /* NucVolume entries in file order      */
    int      n_volumes, cap_volumes;
} mmca_run_cfg;

/* set_defaults (read_ctrl.c:30-108) */
void mmca_ctrl_defaults(mmca_run_cfg *cfg);
/* Reads `path` (and the files it includes) into cfg, which must hold the
 * defaults or an earlier read.  Returns 0, or -1 with a message on stderr. */
int  mmca_ctrl_read(const char *path, mmca_run_cfg *cfg);
void mmca_ctrl_free(mmca_run_cfg *cfg);

/* Parameters of the whole domain (NSubblocks x NCellsPerSB) as one block. */
void mmca_ctrl_params(const mmca_run_cfg *cfg, int check_negative_ce, int precision, mmca_params *p);
/* the subblock of one MPI rank (NCellsPerSB at its Cart coordinates) */
void mmca_ctrl_params_rank(const mmca_run_cfg *cfg, int check_negative_ce, int precision, int rank,
                           mmca_params *p);
/* Nucleation set-up of that block (random stream of rank 0); the lists
 * point into cfg. */
void mmca_ctrl_nucleation(const mmca_run_cfg *cfg, mmca_nucleation_cfg *n);

#endif
