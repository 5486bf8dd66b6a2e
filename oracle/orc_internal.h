/* orc_internal.h -- oracle-private declarations (TEST INFRASTRUCTURE ONLY). */
#pragma once
/* a point (x + sx, y + sy): canonical coordinates plus an integer lattice shift */
typedef struct { double x, y; int sx, sy; } orc_pt;
int orc_orient_pt(const orc_pt* a, const orc_pt* b, const orc_pt* c);
int orc_incircle_pt(const orc_pt* a, const orc_pt* b, const orc_pt* c, const orc_pt* d);
int orc_incircle_sos_pt(const orc_pt* a, const orc_pt* b, const orc_pt* c, const orc_pt* d);
int orc_minimg_shift(double b, double a);
