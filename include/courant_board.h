/* courant_board.h: board model -> simulation front-end of the MI355X courant engine.
 *
 * The board path of the reference's web front-end, as C functions of the same
 * library (libcourant.so), so that batched and served boards need no JavaScript:
 *
 *   board model and KiCad import   web/src/import/kicad.ts:51-220 (Board, importKicad)
 *   copper geometry                web/src/import/kicad.ts:277-387 (copperAt, copperRow,
 *                                  inRow, copperEdges)
 *   rule-based mesher              web/src/mesh/mesher.ts:47-156 (lambdaOver20, meshAxis,
 *                                  withPml, estimate)
 *   board -> simulation            web/src/import/boardsim.ts:70-331 (planBoard,
 *                                  buildBoardSim and their helpers)
 *
 * Lengths in metres; the board's y points up (KiCad's y is flipped on import), z up
 * from the bottom copper.  Errors: -1 or NULL with the message in cem_last_error().
 */
#ifndef COURANT_BOARD_H
#define COURANT_BOARD_H

#include <stddef.h>

#include "courant.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cem_board cem_board;

enum { CEM_LAYER_COPPER = 0, CEM_LAYER_DIELECTRIC = 1 };
enum { CEM_PORT_XLO = 0, CEM_PORT_XHI = 1 };

/* ---- board model (kicad.ts:51-75, 86-220; load.ts:20-29) ---- */
CEM_API cem_board *cem_board_create(void);
/* KiCad 6-8 .kicad_pcb text: stackup, nets, segments, arcs, vias, zone fills,
 * footprint pads (rect, roundrect, oval, circle, custom as its convex hull,
 * through-hole barrels) and the Edge.Cuts bounding box (kicad.ts:86-219). */
CEM_API cem_board *cem_board_import_kicad(const char *text);
CEM_API void cem_board_destroy(cem_board *b);
CEM_API cem_board *cem_board_clone(const cem_board *b);
/* Appends a layer on top of the stack (z is recomputed bottom-up: dielectrics
 * stack, copper is a sheet at the surface, kicad.ts:130-135).  eps_r,
 * loss_tangent, roughness: NaN when not given. */
CEM_API int cem_board_add_layer(cem_board *b, const char *name, int type, double thickness, double eps_r,
                                double loss_tangent, double roughness, int mask);
CEM_API int cem_board_clear_layers(cem_board *b);
CEM_API int cem_board_add_track(cem_board *b, double ax, double ay, double bx, double by, double width,
                                const char *layer, int net);
CEM_API int cem_board_add_via(cem_board *b, double x, double y, double drill, double size, const char *layer_a,
                              const char *layer_b, int net);
CEM_API int cem_board_add_poly(cem_board *b, int npts, const double *xy, const char *layer, int net);
CEM_API int cem_board_set_net(cem_board *b, int net, const char *name);
CEM_API int cem_board_set_outline(cem_board *b, double x0, double y0, double x1, double y1);

/* what: 0 layers, 1 tracks, 2 vias, 3 polygons, 4 nets */
CEM_API int cem_board_count(const cem_board *b, int what);
/* v: thickness, eps_r, loss_tangent, roughness, z0, z1 (NaN where not given) */
CEM_API int cem_board_layer(const cem_board *b, int i, char *name, int cap, int *type, double *v, int *mask);
/* v: ax, ay, bx, by, width */
CEM_API int cem_board_track(const cem_board *b, int i, double *v, char *layer, int cap, int *net);
/* v: x, y, drill, size */
CEM_API int cem_board_via(const cem_board *b, int i, double *v, char *layer_a, char *layer_b, int cap, int *net);
/* returns the vertex count; writes min(count, xy_cap) (x, y) pairs */
CEM_API int cem_board_poly(const cem_board *b, int i, double *xy, int xy_cap, char *layer, int cap, int *net);
CEM_API int cem_board_net(const cem_board *b, int i, int *number, char *name, int cap);
/* x0, y0, x1, y1 */
CEM_API int cem_board_outline(const cem_board *b, double *v);

/* ---- copper geometry (kicad.ts:277-387) ---- */
CEM_API int cem_board_copper_at(const cem_board *b, const char *layer, double x, double y);
/* the row rasteriser of the row y = const evaluated at n positions (same answers as
 * cem_board_copper_at on that row) */
CEM_API int cem_board_copper_row(const cem_board *b, const char *layer, double y, int n, const double *x,
                                 unsigned char *out);
/* axis-aligned copper boundaries of constant coordinate `axis`, sorted, with the
 * metal side (metal_hi 1: metal above x); returns the count, writes min(count, cap) */
CEM_API int cem_board_copper_edges(const cem_board *b, const char *layer, int axis, double *x, int *metal_hi,
                                   int cap);

/* ---- rule-based mesher (mesher.ts:47-156) ---- */
typedef struct {
    double lo, hi;                                  /* domain extent */
    int nsheet; const double *sheet_x;              /* zero-thickness metal edges ... */
    const int *sheet_metal_hi;                      /* ... and the metal side (1: above x) */
    int nplanes; const double *planes;              /* lines exactly on the grid */
    int nspans; const double *spans;                /* (a, b, n) triples: >= n cells across */
    int ncaps; const double *caps;                  /* (a, b, h) triples: maximum cell in [a, b] */
    int nrefine; const double *refine;              /* smallest cells here; -1: every plane and sheet edge */
    int nextra; const double *extra;                /* manual extra lines */
} cem_axis_features;

CEM_API double cem_lambda_over_20(double fmax, double er);
/* Node lines of one axis; returns their count and writes min(count, cap) of them.
 * report: cells, min cell, max cell, max neighbour ratio, warnings; warnings:
 * the messages, one per line (NULL / 0 to skip). ratio <= 0: 1.3. */
CEM_API int cem_mesh_axis(const cem_axis_features *f, double hmin, double hmax, double ratio, double *nodes,
                          int cap, double *report, char *warnings, int wcap);
/* n nodes -> the nodes with npml cells of h added below (lo) and above (hi) */
CEM_API int cem_mesh_with_pml(const double *nodes, int n, int npml, double h, int lo, int hi, double *out);
/* out: cells, points, dt, steps, bytes, cell-steps of a run of t_sim seconds */
CEM_API int cem_mesh_estimate(const double *x, int nx1, const double *y, int ny1, const double *z, int nz1,
                              double t_sim, int gpu, int npml, double *out);

/* ---- board -> simulation (boardsim.ts:22-59, 110-331) ---- */
typedef struct {
    double fmax, df, hmin;        /* band (Hz) and the cell at copper edges (m) */
    double feed, monitor;         /* NaN: 4 mm extrusion past each port edge, 2.5 mm source-monitor */
    double air_above, air_side;   /* NaN: 1.5 mm, 1 mm */
    int npml;                     /* < 0: 8 */
    int lossy;                    /* Djordjevic-Sarkar dielectrics from the stackup + copper loss */
    int thick_copper;             /* copper as PEC bars of its stackup thickness */
    double y_window;              /* > 0: only a band of this half-width around the port tracks */
    int excite;                   /* index of the excited port */
} cem_board_options;

typedef struct {
    const char *net_name;         /* the port track's net by name, or NULL ... */
    int net;                      /* ... by number */
    int edge;                     /* CEM_PORT_XLO / CEM_PORT_XHI */
    double offset;                /* port plane, measured into the board from its edge */
} cem_board_port;

/* The mesh of a board run without building it (the pre-run estimate): cells[3],
 * and the node lines when x / y / z are not NULL (cells[a] + 1 each). */
CEM_API int cem_board_plan(const cem_board *b, const cem_board_options *o, const cem_board_port *ports, int nports,
                           int *cells, double *x, double *y, double *z);
/* The finalized simulation: mode ports at the port planes (port_ids[nports]),
 * the excitation and frequencies set; t_stop = 10.5 tau of the drive. */
CEM_API cem_sim *cem_board_build(const cem_board *b, const cem_board_options *o, const cem_board_port *ports,
                                 int nports, int *port_ids, double *t_stop);

#ifdef __cplusplus
}
#endif

#endif
