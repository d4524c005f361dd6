/* A C caller of the board front-end: imports a .kicad_pcb, plans and builds the board run
 * of its RF net between the board edges, and prints the mesh and port ids. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "courant_board.h"

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *text = malloc((size_t)n + 1);
    if (!text || fread(text, 1, (size_t)n, f) != (size_t)n) return 2;
    text[n] = 0;
    fclose(f);
    cem_board *b = cem_board_import_kicad(text);
    free(text);
    if (!b) {
        printf("error %s\n", cem_last_error());
        return 1;
    }
    cem_board_options o = {40e9, 1e9, 40e-6, NAN, NAN, NAN, NAN, -1, 0, 0, 0.0, 0};
    cem_board_port ports[2] = {{"RF", 0, CEM_PORT_XLO, 0.0}, {"RF", 0, CEM_PORT_XHI, 0.0}};
    int cells[3];
    if (cem_board_plan(b, &o, ports, 2, cells, NULL, NULL, NULL) != 0) {
        printf("error %s\n", cem_last_error());
        return 1;
    }
    int ids[2];
    double t_stop;
    cem_sim *s = cem_board_build(b, &o, ports, 2, ids, &t_stop);
    if (!s) {
        printf("error %s\n", cem_last_error());
        return 1;
    }
    int nx, ny, nz;
    cem_dims(s, &nx, &ny, &nz);
    printf("plan %d %d %d sim %d %d %d ports %d %d t_stop %.17g nfreq %d\n", cells[0], cells[1], cells[2], nx, ny, nz,
           ids[0], ids[1], t_stop, cem_nfreq(s));
    cem_destroy(s);
    cem_board_destroy(b);
    return 0;
}
