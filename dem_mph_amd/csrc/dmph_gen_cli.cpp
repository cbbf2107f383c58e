// GeneratorForMph: the reference's scene generator program
// (generator/generator.cpp:117-127).  `GeneratorForMph [name]` reads
// name.boid and writes name.grid (default sample.boid -> sample.grid).  As the
// generator does, the particles read before a parse error are still written
// and the exit status stays 0 for parse errors; a .grid that cannot be written
// exits 1.
#include <cstdio>
#include <string>

#include "dem_mph.h"

int main(int argc, char **argv) {
    std::string boid = "sample.boid", grid = "sample.grid";
    if (argc > 1) {
        boid = std::string(argv[1]) + ".boid";
        grid = std::string(argv[1]) + ".grid";
    }
    dmph_grid g;
    const int rc = dmph_boid_generate(boid.c_str(), &g);
    if (rc != DMPH_OK) fprintf(stderr, "error: \n\tfile:%s\n\t%s\n", boid.c_str(), dmph_boid_error());
    fprintf(stderr, "%d particles were generated\n", g.n);
    const int w = dmph_grid_write(grid.c_str(), &g);
    dmph_grid_free(&g);
    if (w != DMPH_OK) {
        fprintf(stderr, "cannot write %s\n", grid.c_str());
        return 1;
    }
    return 0;
}
