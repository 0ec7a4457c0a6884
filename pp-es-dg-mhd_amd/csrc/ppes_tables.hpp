// Host-side set-up tables (see ppes_tables.cpp).
#pragma once

namespace ppes {
// GLL nodes/weights, differentiation and Legendre maps for degree p (0..6;
// degree 0 is the one-node rule); returns 0 on success. Null pointers are skipped.
int host_tables(int p, double *x, double *w, double *D, double *Pmn, double *Pnm);
// LDF projector in Legendre modal coordinates, 2N^2 x 2N^2 column-major
int host_ldf_modal(int p, double *op);
// LDF projection of degrees (px, py) as a dense nodal operator on
// (Bx - avg, By - avg), 2NN x 2NN with NN = (px+1)(py+1)
int host_ldf_nodal(int px, int py, double hx, double hy, double *op);
}  // namespace ppes
