// Scene generator: .boid shape descriptions -> a particle set in .grid terms
// (generator/generator.cpp).  Host code; the particle counts of a scene are
// small next to a step, so nothing here runs on the GPU.
//
// File syntax (generator.cpp:129-183, 187-658): header lines `ParticleDistance
// d`, `LowerDomain x y z`, `UpperDomain x y z` ('#' lines skipped); blocks
// `StartX ... EndX` for X in Cuboid, Cuboid2, Cyboid, Cyboid2, Recboid,
// Recboid2, each a whitespace-separated list of `Key values` pairs.  Shapes are
// generated kind by kind in that order (each kind in file order), not in file
// order (generator.cpp:660-944).  The lattice loops keep the generator's
// arithmetic -- accumulated coordinates, the same bounds, the same rotations
// (Cuboid2 takes Angle as radians; Recboid and Recboid2 take degrees through
// 3.1415 / 180) -- so the written file matches the generator's byte for byte.
//
// Deviations, all on malformed input: a block that reaches end of file without
// its End token is an error (the generator loops forever); a file that cannot
// be opened is an error (the generator writes an empty particle set).  In both
// cases, and on the generator's own errors (unknown key, missing required key,
// unreadable value), the shapes read before the error are still generated, as
// the generator does, and the call returns DMPH_E_FORMAT / DMPH_E_IO.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dem_mph.h"

namespace {

enum Kind { CUBOID = 0, CUBOID2, CYBOID, CYBOID2, RECBOID, RECBOID2, NKIND };
const char *const kName[NKIND] = {"Cuboid", "Cuboid2", "Cyboid", "Cyboid2", "Recboid", "Recboid2"};

struct Shape {
    double space = 0.0;
    int type = 0;
    double lower[3] = {0, 0, 0}, upper[3] = {0, 0, 0}, vel[3] = {0, 0, 0};
    double angle = 0.0, ratio = 0.0;
};

// keys each kind accepts; bit order: Spacing Type Lower Upper Velocity Enthalpy Angle Ratio
enum { K_SPACING = 1, K_TYPE = 2, K_LOWER = 4, K_UPPER = 8, K_VEL = 16, K_ENTH = 32, K_ANGLE = 64, K_RATIO = 128 };
const unsigned kAccept[NKIND] = {
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL | K_ENTH,                 // Cuboid (:199-235)
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL | K_ANGLE,                // Cuboid2 (:276-314)
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL | K_ENTH | K_RATIO,        // Cyboid (:355-399)
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL | K_ENTH | K_RATIO,        // Cyboid2 (:437-480)
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL | K_ENTH | K_ANGLE,        // Recboid (:515-558)
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL | K_ENTH | K_ANGLE,        // Recboid2 (:593-636)
};
const unsigned kRequire[NKIND] = {
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL,
    K_SPACING | K_TYPE | K_LOWER | K_UPPER | K_VEL | K_ANGLE,
    kAccept[CYBOID], kAccept[CYBOID2], kAccept[RECBOID], kAccept[RECBOID2],
};

struct Out {
    std::vector<int> type;
    std::vector<double> pos, vel;
    void add(int t, double x, double y, double z, const double v[3]) {
        type.push_back(t);
        pos.push_back(x); pos.push_back(y); pos.push_back(z);
        vel.push_back(v[0]); vel.push_back(v[1]); vel.push_back(v[2]);
    }
};

// Reads one block's `Key values` pairs up to its End token.  0 = ok, 1 = the
// generator's errors (unknown key, unreadable value, missing required key),
// 2 = end of file inside the block.
int read_block(FILE *fp, Kind kind, Shape &s, std::string &msg) {
    const std::string end = std::string("End") + kName[kind];
    unsigned seen = 0;
    char tok[256];
    double enthalpy = 0.0;
    for (;;) {
        if (fscanf(fp, "%255s", tok) != 1) {
            msg = "end of file inside Start" + std::string(kName[kind]) + " (no " + end + ")";
            return 2;
        }
        if (end == tok) break;
        unsigned key = 0;
        int ok = 1;
        if (!strcmp(tok, "Spacing")) { key = K_SPACING; ok = fscanf(fp, "%lf", &s.space) == 1; }
        else if (!strcmp(tok, "Type")) { key = K_TYPE; ok = fscanf(fp, "%d", &s.type) == 1; }
        else if (!strcmp(tok, "Lower")) { key = K_LOWER; for (int k = 0; k < 3 && ok; k++) ok = fscanf(fp, "%lf", &s.lower[k]) == 1; }
        else if (!strcmp(tok, "Upper")) { key = K_UPPER; for (int k = 0; k < 3 && ok; k++) ok = fscanf(fp, "%lf", &s.upper[k]) == 1; }
        else if (!strcmp(tok, "Velocity")) { key = K_VEL; for (int k = 0; k < 3 && ok; k++) ok = fscanf(fp, "%lf", &s.vel[k]) == 1; }
        else if (!strcmp(tok, "Enthalpy")) { key = K_ENTH; ok = fscanf(fp, "%lf", &enthalpy) == 1; }
        else if (!strcmp(tok, "Angle")) { key = K_ANGLE; ok = fscanf(fp, "%lf", &s.angle) == 1; }
        else if (!strcmp(tok, "Ratio")) { key = K_RATIO; ok = fscanf(fp, "%lf", &s.ratio) == 1; }
        if (!(key & kAccept[kind])) {
            msg = std::string("no such indication: ") + tok + " in Start" + kName[kind];
            return 1;
        }
        if (!ok) {
            msg = std::string("unreadable value of ") + tok + " in Start" + kName[kind];
            return 1;
        }
        seen |= key;
    }
    if ((seen & kRequire[kind]) != kRequire[kind]) {
        msg = std::string("Start") + kName[kind] + " misses a required key";
        return 1;
    }
    return 0;
}

// Lattice counts and spacings of a box (generator.cpp:665-666): round(width /
// space) per axis, then width / count.
void box_spacing(const Shape &s, double w[3], double sp[3]) {
    for (int k = 0; k < 3; k++) {
        w[k] = s.upper[k] - s.lower[k];
        const int c = (int)round(w[k] / s.space);
        sp[k] = w[k] / c;
    }
}

// Cuboid (:663-685): cell centres, bounds at +0.5 and -0.49 spacings
void gen_cuboid(const Shape &s, Out &o) {
    double w[3], sp[3];
    box_spacing(s, w, sp);
    for (double x = s.lower[0] + 0.5 * sp[0]; x < s.upper[0] - 0.49 * sp[0]; x += sp[0])
        for (double y = s.lower[1] + 0.5 * sp[1]; y < s.upper[1] - 0.49 * sp[1]; y += sp[1])
            for (double z = s.lower[2] + 0.5 * sp[2]; z < s.upper[2] - 0.49 * sp[2]; z += sp[2]) o.add(s.type, x, y, z, s.vel);
}

// Cuboid2 (:687-722): x, y from 0.01 spacing in, rotated by Angle (radians)
// about the box centre in the xy plane
void gen_cuboid2(const Shape &s, Out &o) {
    double w[3], sp[3];
    box_spacing(s, w, sp);
    double c[3];
    for (int k = 0; k < 3; k++) c[k] = (s.upper[k] + s.lower[k]) / 2;
    const double ca = cos(s.angle), sa = sin(s.angle);
    for (double x = s.lower[0] + 0.01 * sp[0]; x < s.upper[0] - 0.0 * sp[0]; x += sp[0])
        for (double y = s.lower[1] + 0.01 * sp[1]; y < s.upper[1] - 0.0 * sp[1]; y += sp[1])
            for (double z = s.lower[2] + 0.5 * sp[2]; z < s.upper[2] - 0.49 * sp[2]; z += sp[2]) {
                const double nx = (x - c[0]) * ca - (y - c[1]) * sa + c[0];
                const double ny = (x - c[0]) * sa + (y - c[1]) * ca + c[1];
                o.add(s.type, nx, ny, z, s.vel);
            }
}

// Cyboid (:724-755): a regular grid whose step is width / floor(width / side),
// side = min width / sqrt 3, points at cell centres (Spacing and Ratio unused)
void gen_cyboid(const Shape &s, Out &o) {
    double w[3];
    for (int k = 0; k < 3; k++) w[k] = s.upper[k] - s.lower[k];
    const double side = fmin(w[0], fmin(w[1], w[2])) / sqrt(3.0);
    int n[3];
    double sp[3];
    for (int k = 0; k < 3; k++) n[k] = (int)floor(w[k] / side);
    for (int k = 0; k < 3; k++) sp[k] = w[k] / n[k];
    for (int i = 0; i < n[0]; ++i)
        for (int j = 0; j < n[1]; ++j)
            for (int k = 0; k < n[2]; ++k)
                o.add(s.type, s.lower[0] + (i + 0.5) * sp[0], s.lower[1] + (j + 0.5) * sp[1],
                      s.lower[2] + (k + 0.5) * sp[2], s.vel);
}

// one ring of a Cyboid2 (:812-826): n = max(6, (int)(2 pi r / spacing)) points
void ring(const Shape &s, const double c[3], double r, double z, Out &o) {
    const double circ = 2.0 * M_PI * r;
    const int n = std::max(6, (int)(circ / s.space));
    for (int t = 0; t < n; t++) {
        const double th = (2.0 * M_PI * t) / n;
        o.add(s.type, c[0] + r * cos(th), c[1] + r * sin(th), z, s.vel);
    }
}

// Cyboid2 (:795-862): the lateral surface of a truncated cone, radius R1 =
// (wx + wy) / 4 at the top and R1 Ratio at the bottom, in rings one spacing
// apart, then the top and the bottom circles
void gen_cyboid2(const Shape &s, Out &o) {
    double w[3], c[3];
    for (int k = 0; k < 3; k++) {
        w[k] = s.upper[k] - s.lower[k];
        c[k] = 0.5 * (s.upper[k] + s.lower[k]);
    }
    const double h = w[2];
    const double r1 = 0.5 * (w[0] + w[1]) * 0.5, r2 = r1 * s.ratio;
    for (double z = s.lower[2] + 0.5 * s.space; z < s.upper[2] - 0.49 * s.space; z += s.space)
        ring(s, c, r2 + (r1 - r2) * ((z - s.lower[2]) / h), z, o);
    ring(s, c, r1, s.upper[2], o);
    ring(s, c, r2, s.lower[2], o);
}

// Recboid (:866-899): the Cuboid2 lattice without rotation, kept where
// tan(Angle deg) > y / x; Recboid2 (:901-934): the same lattice rotated by Angle
// degrees about the origin.  The y bound uses the x spacing (:876, :910).
void gen_recboid(const Shape &s, bool rotate, Out &o) {
    double w[3], sp[3];
    box_spacing(s, w, sp);
    for (double x = s.lower[0] + 0.01 * sp[0]; x < s.upper[0] - 0.0 * sp[0]; x += sp[0])
        for (double y = s.lower[1] + 0.01 * sp[1]; y < s.upper[1] - 0.0 * sp[0]; y += sp[1])
            for (double z = s.lower[2] + 0.5 * sp[2]; z < s.upper[2] - 0.49 * sp[2]; z += sp[2]) {
                if (!rotate) {
                    if (tan(s.angle * 3.1415 / 180) > y / x) o.add(s.type, x, y, z, s.vel);
                } else {
                    const double a = s.angle * 3.1415 / 180;
                    o.add(s.type, x * cos(a) - y * sin(a), x * sin(a) + y * cos(a), z, s.vel);
                }
            }
}

thread_local std::string g_err;

}  // namespace

extern "C" {

int dmph_boid_generate(const char *path, dmph_grid *out) {
    if (!path || !out) return DMPH_E_ARG;
    memset(out, 0, sizeof *out);
    out->spacing = -1.0;   // the generator's value when ParticleDistance is missing
    std::vector<Shape> shapes[NKIND];
    int rc = DMPH_OK;
    g_err.clear();
    FILE *fp = fopen(path, "r");
    if (!fp) {
        g_err = std::string("cannot open ") + path;
        rc = DMPH_E_IO;
    }
    // header lines and blocks (generator.cpp:129-183)
    char buf[1024], tok[256];
    while (fp && rc == DMPH_OK && fgets(buf, sizeof buf, fp)) {
        if (buf[0] == '#') continue;
        if (sscanf(buf, "%255s", tok) != 1) continue;
        int ok = 1;
        if (!strcmp(tok, "ParticleDistance")) ok = sscanf(buf, " %*s %lf", &out->spacing) == 1;
        else if (!strcmp(tok, "LowerDomain"))
            ok = sscanf(buf, " %*s %lf %lf %lf", &out->domain_min[0], &out->domain_min[1], &out->domain_min[2]) == 3;
        else if (!strcmp(tok, "UpperDomain"))
            ok = sscanf(buf, " %*s %lf %lf %lf", &out->domain_max[0], &out->domain_max[1], &out->domain_max[2]) == 3;
        else
            for (int k = 0; k < NKIND; k++)
                if (!strcmp(tok, (std::string("Start") + kName[k]).c_str())) {
                    Shape s;
                    const int r = read_block(fp, (Kind)k, s, g_err);
                    if (r == 0) shapes[k].push_back(s);
                    else rc = DMPH_E_FORMAT;
                    break;
                }
        if (!ok) {
            g_err = std::string("unreadable header line: ") + buf;
            rc = DMPH_E_FORMAT;
        }
    }
    if (fp) fclose(fp);
    // generation, kind by kind (generator.cpp:660-944)
    Out o;
    for (const Shape &s : shapes[CUBOID]) gen_cuboid(s, o);
    for (const Shape &s : shapes[CUBOID2]) gen_cuboid2(s, o);
    for (const Shape &s : shapes[CYBOID]) gen_cyboid(s, o);
    for (const Shape &s : shapes[CYBOID2]) gen_cyboid2(s, o);
    for (const Shape &s : shapes[RECBOID]) gen_recboid(s, false, o);
    for (const Shape &s : shapes[RECBOID2]) gen_recboid(s, true, o);
    const size_t n = o.type.size();
    out->n = (int)n;
    out->time = 0.0;
    out->type = (int *)malloc(std::max<size_t>(n, 1) * sizeof(int));
    out->pos = (double *)malloc(std::max<size_t>(n, 1) * 3 * sizeof(double));
    out->vel = (double *)malloc(std::max<size_t>(n, 1) * 3 * sizeof(double));
    if (!out->type || !out->pos || !out->vel) {
        dmph_grid_free(out);
        g_err = "out of host memory";
        return DMPH_E_ALLOC;
    }
    if (n) {
        memcpy(out->type, o.type.data(), n * sizeof(int));
        memcpy(out->pos, o.pos.data(), 3 * n * sizeof(double));
        memcpy(out->vel, o.vel.data(), 3 * n * sizeof(double));
    }
    return rc;
}

const char *dmph_boid_error(void) { return g_err.c_str(); }

int dmph_grid_write(const char *path, const dmph_grid *g) {
    if (!path || !g || g->n < 0) return DMPH_E_ARG;
    FILE *fp = fopen(path, "w");
    if (!fp) return DMPH_E_IO;
    // the generator's layout (generator.cpp:948-974)
    fprintf(fp, "%lf\n", g->time);
    fprintf(fp, "%d %e  %e %e %e  %e %e %e\n", g->n, g->spacing, g->domain_min[0], g->domain_max[0], g->domain_min[1],
            g->domain_max[1], g->domain_min[2], g->domain_max[2]);
    for (int i = 0; i < g->n; i++)
        fprintf(fp, "%d    %e %e %e  %e %e %e \n", g->type[i], g->pos[3 * i], g->pos[3 * i + 1], g->pos[3 * i + 2],
                g->vel[3 * i], g->vel[3 * i + 1], g->vel[3 * i + 2]);
    const int bad = ferror(fp);
    return fclose(fp) != 0 || bad ? DMPH_E_IO : DMPH_OK;
}

}  // extern "C"
