/*
 * Control-file reader of the C host driver.  Format (read_ctrl.c:235-812):
 * one "Key value..." per line, keys case-insensitive, lines whose first
 * token starts with '#' or '%' are comments, values are read with sscanf so
 * trailing text is ignored, unknown keys are ignored (read_ctrl.c:794-800),
 * GeoFileName / MatFileName / AlloyPropsFile0 include another file
 * (read_ctrl.c:623-631, relative to the including file's directory here).
 */
#define _GNU_SOURCE
#include "mmca_ctrl.h"

#include <libgen.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>

void
mmca_ctrl_defaults(mmca_run_cfg *c)
{
    /* set_defaults (read_ctrl.c:30-108); float literals as the reference stores them */
    memset(c, 0, sizeof(*c));
    snprintf(c->basefilename, sizeof(c->basefilename), "output");
    c->pad[0] = c->pad[1] = c->pad[2] = 1;
    c->nsub[0] = c->nsub[1] = c->nsub[2] = 1;
    c->cell_size = 0.01f;
    c->courant_limit = 0.2;
    c->t_init = 100.0f;
    c->cool_rate = 0.0f;
    c->liquidus_temp = 1180.0f;
    c->solidus_temp = 1120.0f;
    c->random_seed = (uint32_t) time(NULL);
    c->fs_finish = 1.0f;
    c->finish_time = 0.0f;
    c->dt = 0.1f;
    c->screenpfreq = 1;
    c->max_total_grains = 1000000;
    c->do_surface = 1;
    c->m_solute[0] = 0.0;
    c->m_solute[1] = 0.0;
    c->m_solute[2] = -2.0;
    c->c_init = 5.0;
    c->d_sol = 1e-13;
    c->d_liq = 1e-9;
    c->part_coef = 0.3;
    c->am_layers = 1;
    c->am_tracks = 1;
    c->vis_format = MMCA_VIS_NONE;
    c->randomize_grains = 1;
}

void
mmca_ctrl_free(mmca_run_cfg *c)
{
    free(c->fixed);
    free(c->volumes);
    c->fixed = NULL;
    c->volumes = NULL;
    c->n_fixed = c->cap_fixed = c->n_volumes = c->cap_volumes = 0;
}

static int
need(int got, int want, const char *key, const char *path)
{
    if (got == want)
        return 0;
    fprintf(stderr, "%s: cannot parse the value of %s\n", path, key);
    return -1;
}

static int
apply(mmca_run_cfg *c, const char *key, const char *rest, const char *path)
{
    double f[8];
    int v[6];
    long long l;
    unsigned long long u;
    if (!strcasecmp(key, "BaseFileName"))
    {
        char name[256];
        if (need(sscanf(rest, "%255s", name), 1, key, path))
            return -1;
        snprintf(c->basefilename, sizeof(c->basefilename), "%s", name);
    }
    else if (!strcasecmp(key, "NSubblocks"))
        return need(sscanf(rest, "%d %d %d", &c->nsub[0], &c->nsub[1], &c->nsub[2]), 3, key, path);
    else if (!strcasecmp(key, "NCellsPerSB"))
        return need(sscanf(rest, "%d %d %d", &c->sbdim[0], &c->sbdim[1], &c->sbdim[2]), 3, key, path);
    else if (!strcasecmp(key, "CellSize"))
        return need(sscanf(rest, "%lf", &c->cell_size), 1, key, path);
    else if (!strcasecmp(key, "CourantLimit"))
        return need(sscanf(rest, "%lf", &c->courant_limit), 1, key, path);
    else if (!strcasecmp(key, "FaceCtrl"))
    {
        if (need(sscanf(rest, "%d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]), 6, key, path))
            return -1;
        if (v[0] != v[1] || v[2] != v[3] || v[4] != v[5])
        {
            fprintf(stderr, "%s: FaceCtrl: both faces of an axis must match (read_ctrl.c:293-296)\n", path);
            return -1;
        }
        c->pad[0] = v[0];
        c->pad[1] = v[2];
        c->pad[2] = v[4];
    }
    else if (!strcasecmp(key, "FixedNuc"))
    {
        if (need(sscanf(rest, "%lf %lf %lf %lf %lf %lf %lf", &f[0], &f[1], &f[2], &f[3], &f[4], &f[5], &f[6]), 7,
                 key, path))
            return -1;
        if (c->n_fixed == c->cap_fixed)
        {
            c->cap_fixed = c->cap_fixed ? 2 * c->cap_fixed : 8;
            c->fixed = realloc(c->fixed, sizeof(*c->fixed) * (size_t) c->cap_fixed);
        }
        mmca_fixed_nuc *n = &c->fixed[c->n_fixed++];
        n->x = f[0];
        n->y = f[1];
        n->z = f[2];
        n->ax = f[3];
        n->ay = f[4];
        n->az = f[5];
        n->threshold = f[6];
    }
    else if (!strcasecmp(key, "NucVolume"))
    {
        if (need(sscanf(rest, "%lf %lf %lf %lf %lf %lf %lf %lf", &f[0], &f[1], &f[2], &f[3], &f[4], &f[5], &f[6],
                        &f[7]), 8, key, path))
            return -1;
        if (c->n_volumes == c->cap_volumes)
        {
            c->cap_volumes = c->cap_volumes ? 2 * c->cap_volumes : 8;
            c->volumes = realloc(c->volumes, sizeof(*c->volumes) * (size_t) c->cap_volumes);
        }
        mmca_nuc_volume *n = &c->volumes[c->n_volumes++];
        n->density = f[0] / 100.0;     /* read_ctrl.c:318: percent */
        n->minx = f[1];
        n->miny = f[2];
        n->minz = f[3];
        n->maxx = f[4];
        n->maxy = f[5];
        n->maxz = f[6];
        n->threshold = f[7];
    }
    else if (!strcasecmp(key, "MaxGrainDensity"))
        return need(sscanf(rest, "%lf", &c->bulk_density), 1, key, path);
    else if (!strcasecmp(key, "MaxGrainDensitySurf"))
        return need(sscanf(rest, "%lf", &c->surf_density), 1, key, path);
    else if (!strcasecmp(key, "MaxTotGrains"))
    {
        if (need(sscanf(rest, "%lld", &l), 1, key, path))
            return -1;
        c->max_total_grains = l;
    }
    else if (!strcasecmp(key, "MouldNuc") || !strcasecmp(key, "MoldNuc"))
        return need(sscanf(rest, "%d", &c->do_surface), 1, key, path);
    else if (!strcasecmp(key, "GNGaussCentre") || !strcasecmp(key, "GNGaussCenter"))
        return need(sscanf(rest, "%lf", &c->bulk_mean), 1, key, path);
    else if (!strcasecmp(key, "GNGaussSigma"))
        return need(sscanf(rest, "%lf", &c->bulk_sigma), 1, key, path);
    else if (!strcasecmp(key, "GNGaussCentreSurf") || !strcasecmp(key, "GNGaussCenterSurf"))
        return need(sscanf(rest, "%lf", &c->surf_mean), 1, key, path);
    else if (!strcasecmp(key, "GNGaussSigmaSurf"))
        return need(sscanf(rest, "%lf", &c->surf_sigma), 1, key, path);
    else if (!strcasecmp(key, "GNOriented"))
        return need(sscanf(rest, "%d", &c->oriented), 1, key, path);
    else if (!strcasecmp(key, "LiquidusTemp"))
        return need(sscanf(rest, "%lf", &c->liquidus_temp), 1, key, path);
    else if (!strcasecmp(key, "SolidusTemp"))
        return need(sscanf(rest, "%lf", &c->solidus_temp), 1, key, path);
    else if (!strcasecmp(key, "Cinit"))
        return need(sscanf(rest, "%lf", &c->c_init), 1, key, path);
    else if (!strcasecmp(key, "m_solute0"))
        return need(sscanf(rest, "%lf %lf %lf", &c->m_solute[0], &c->m_solute[1], &c->m_solute[2]), 3, key, path);
    else if (!strcasecmp(key, "part_coef0"))
    {
        /* k = a T^2 + b T + c; the step uses c (read_ctrl.c:404-417) */
        if (need(sscanf(rest, "%lf %lf %lf", &f[0], &f[1], &f[2]), 3, key, path))
            return -1;
        c->part_coef = f[2];
    }
    else if (!strcasecmp(key, "Dsol"))
        return need(sscanf(rest, "%lf", &c->d_sol), 1, key, path);
    else if (!strcasecmp(key, "Dliq"))
        return need(sscanf(rest, "%lf", &c->d_liq), 1, key, path);
    else if (!strcasecmp(key, "InitialTemperature"))
        return need(sscanf(rest, "%lf", &c->t_init), 1, key, path);
    else if (!strcasecmp(key, "Velo_Coef"))
        return need(sscanf(rest, "%lf", &c->velo_coef), 1, key, path);
    else if (!strcasecmp(key, "Grad_Coef"))
        return need(sscanf(rest, "%lf", &c->grad_coef), 1, key, path);
    else if (!strcasecmp(key, "Gradient"))
        return need(sscanf(rest, "%lf", &c->gradient), 1, key, path);
    else if (!strcasecmp(key, "Velocity"))
        return need(sscanf(rest, "%lf", &c->velocity), 1, key, path);
    else if (!strcasecmp(key, "CoolingRate"))
        return need(sscanf(rest, "%lf", &c->cool_rate), 1, key, path);
    else if (!strcasecmp(key, "FSGrow"))
        return need(sscanf(rest, "%lf", &c->fsgrow), 1, key, path);
    else if (!strcasecmp(key, "TimeStep"))
        return need(sscanf(rest, "%lf", &c->dt), 1, key, path);
    else if (!strcasecmp(key, "FinishTime"))
        return need(sscanf(rest, "%lf", &c->finish_time), 1, key, path);
    else if (!strcasecmp(key, "FsFinish"))
        return need(sscanf(rest, "%lf", &c->fs_finish), 1, key, path);
    else if (!strcasecmp(key, "ScreenPFreq"))
        return need(sscanf(rest, "%d", &c->screenpfreq), 1, key, path);
    else if (!strcasecmp(key, "DataWriteFreq"))
        return need(sscanf(rest, "%d", &c->data_write_freq), 1, key, path);
    else if (!strcasecmp(key, "DataWriteStart"))
        return need(sscanf(rest, "%d", &c->data_write_start), 1, key, path);
    else if (!strcasecmp(key, "RandomizeGrains"))
        return need(sscanf(rest, "%d", &c->randomize_grains), 1, key, path);
    else if (!strcasecmp(key, "DataFormat"))
    {
        char fmt[64] = { 0 };
        if (need(sscanf(rest, "%63s", fmt), 1, key, path))
            return -1;
        if (!strcasecmp(fmt, "XDMF"))
            c->vis_format = MMCA_VIS_XDMF;
        else if (!strcasecmp(fmt, "VTK"))
            c->vis_format = MMCA_VIS_VTK;
        else
        {
            /* read_ctrl.c:601-620 takes VTK or XDMF only (NONE is the default) */
            fprintf(stderr, "%s: unsupported data file type: %s\n", path, fmt);
            return -1;
        }
        return 0;
    }
    else if (!strcasecmp(key, "GeoFileName") || !strcasecmp(key, "MatFileName") ||
             !strcasecmp(key, "AlloyPropsFile0"))
    {
        char inc[1024], full[2048];
        if (need(sscanf(rest, "%1023s", inc), 1, key, path))
            return -1;
        if (inc[0] == '/')
            snprintf(full, sizeof(full), "%s", inc);
        else
        {
            char dirbuf[1024];
            snprintf(dirbuf, sizeof(dirbuf), "%s", path);
            snprintf(full, sizeof(full), "%s/%s", dirname(dirbuf), inc);
        }
        return mmca_ctrl_read(full, c);
    }
    else if (!strcasecmp(key, "RandSeedVal"))
    {
        if (need(sscanf(rest, "%llu", &u), 1, key, path))
            return -1;
        c->random_seed = (uint32_t) u;
    }
    else if (!strcasecmp(key, "AnalyticalThermal"))
        return need(sscanf(rest, "%d", &c->thermal_ana), 1, key, path);
    else if (!strcasecmp(key, "LaserPower"))
        return need(sscanf(rest, "%lf", &c->laser_power), 1, key, path);
    else if (!strcasecmp(key, "LaserSpeed"))
        return need(sscanf(rest, "%lf", &c->laser_speed), 1, key, path);
    else if (!strcasecmp(key, "LaserSpotSize"))
        return need(sscanf(rest, "%lf", &c->laser_spot), 1, key, path);
    else if (!strcasecmp(key, "LaserAbsorption"))
        return need(sscanf(rest, "%lf", &c->laser_absorp), 1, key, path);
    else if (!strcasecmp(key, "LaserHatchSpacing"))
        return need(sscanf(rest, "%lf", &c->laser_hatch), 1, key, path);
    else if (!strcasecmp(key, "SubInitialTemp"))
        return need(sscanf(rest, "%lf", &c->sub_init_temp), 1, key, path);
    else if (!strcasecmp(key, "ThermalConductivity"))
        return need(sscanf(rest, "%lf", &c->th_conductivity), 1, key, path);
    else if (!strcasecmp(key, "Density"))
        return need(sscanf(rest, "%lf", &c->density), 1, key, path);
    else if (!strcasecmp(key, "SpecificHeat"))
        return need(sscanf(rest, "%lf", &c->specific_heat), 1, key, path);
    else if (!strcasecmp(key, "LaserStartLocation"))
        return need(sscanf(rest, "%lf %lf %lf", &c->laser_start[0], &c->laser_start[1], &c->laser_start[2]), 3,
                    key, path);
    else if (!strcasecmp(key, "NoOfLayers"))
        return need(sscanf(rest, "%d", &c->am_layers), 1, key, path);
    else if (!strcasecmp(key, "NoOfTracks"))
        return need(sscanf(rest, "%d", &c->am_tracks), 1, key, path);
    /* every other key is accepted and ignored, as in the reference */
    return 0;
}

int
mmca_ctrl_read(const char *path, mmca_run_cfg *c)
{
    FILE *fp = fopen(path, "r");
    if (!fp)
    {
        fprintf(stderr, "cannot open control file %s\n", path);
        return -1;
    }
    char line[4096];
    int rc = 0;
    while (rc == 0 && fgets(line, sizeof(line), fp))
    {
        line[255] = '\0';               /* the reference reads 256-byte lines */
        char *p = line;
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')
            p++;
        if (!*p || *p == '#' || *p == '%')
            continue;
        char *k = p;
        while (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n')
            p++;
        char key[256];
        size_t kl = (size_t) (p - k);
        if (kl >= sizeof(key))
            kl = sizeof(key) - 1;
        memcpy(key, k, kl);
        key[kl] = '\0';
        rc = apply(c, key, p, path);
    }
    fclose(fp);
    return rc;
}

void
mmca_ctrl_params(const mmca_run_cfg *c, int check_negative_ce, int precision, mmca_params *p)
{
    memset(p, 0, sizeof(*p));
    for (int a = 0; a < 3; a++)
    {
        p->gdim[a] = c->nsub[a] * c->sbdim[a];
        p->pad_bndy[a] = c->pad[a];
    }
    p->nx = p->gdim[0];
    p->ny = p->gdim[1];
    p->nz = p->gdim[2];
    p->cell_size = c->cell_size;
    p->dt = c->dt;
    p->courant_limit = c->courant_limit;
    p->c_init = c->c_init;
    p->liquidus_temp = c->liquidus_temp;
    p->solidus_temp = c->solidus_temp;
    for (int a = 0; a < 3; a++)
        p->m_solute[a] = c->m_solute[a];
    p->part_coef = c->part_coef;
    p->d_sol = c->d_sol;
    p->d_liq = c->d_liq;
    p->fsgrow = c->fsgrow;
    p->temp_model = c->thermal_ana ? MMCA_TEMP_AM_BEAM
                  : (c->gradient != 0 ? MMCA_TEMP_GRADIENT : MMCA_TEMP_UNIFORM);
    p->am_layers = c->am_layers;
    p->am_tracks = c->am_tracks;
    p->gnsbx = 1;                       /* one block spans the whole x extent */
    p->t_init = c->t_init;
    p->cool_rate = c->cool_rate;
    p->gradient = c->gradient;
    p->grad_coef = c->grad_coef;
    p->velocity = c->velocity;
    p->velo_coef = c->velo_coef;
    p->laser_power = c->laser_power;
    p->laser_speed = c->laser_speed;
    p->laser_spot = c->laser_spot;
    p->laser_absorp = c->laser_absorp;
    p->laser_hatch = c->laser_hatch;
    p->sub_init_temp = c->sub_init_temp;
    p->th_conductivity = c->th_conductivity;
    p->density = c->density;
    p->specific_heat = c->specific_heat;
    for (int a = 0; a < 3; a++)
        p->laser_start[a] = c->laser_start[a];
    p->max_total_grains = c->max_total_grains;
    p->check_negative_ce = check_negative_ce;
    p->precision = precision;
    p->timestep = 0;
}

/* The subblock of MPI rank `rank` (ca_main.c:197-212: Cart grid, x fastest). */
void
mmca_ctrl_params_rank(const mmca_run_cfg *c, int check_negative_ce, int precision, int rank, mmca_params *p)
{
    mmca_ctrl_params(c, check_negative_ce, precision, p);
    p->nx = c->sbdim[0];
    p->ny = c->sbdim[1];
    p->nz = c->sbdim[2];
    p->sb_coords[0] = rank % c->nsub[0];
    p->sb_coords[1] = (rank / c->nsub[0]) % c->nsub[1];
    p->sb_coords[2] = rank / (c->nsub[0] * c->nsub[1]);
    p->gnsbx = c->nsub[0];              /* the beam track spans gnsbx * nx (temperature.c:68) */
}

void
mmca_ctrl_nucleation(const mmca_run_cfg *c, mmca_nucleation_cfg *n)
{
    memset(n, 0, sizeof(*n));
    n->random_seed = c->random_seed;
    n->rank = 0;
    n->oriented = c->oriented;
    n->do_surface = c->do_surface;
    n->bulk_density = c->bulk_density;
    n->bulk_mean = c->bulk_mean;
    n->bulk_sigma = c->bulk_sigma;
    n->surf_density = c->surf_density;
    n->surf_mean = c->surf_mean;
    n->surf_sigma = c->surf_sigma;
    n->n_fixed = c->n_fixed;
    n->fixed = c->fixed;
    n->n_volumes = c->n_volumes;
    n->volumes = c->volumes;
}
