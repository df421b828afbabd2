/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * Cubed-sphere topology, local-mesh vertex ids, 3D connectivity, dofmaps and the
 * cubed-sphere 4-colouring, restated from the reference's Fortran in its own loop order.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

/* direction / corner ids of the reference element
 * (infrastructure/source/mesh/reference_element_mod.F90:336-353) */
enum { W = 1, S = 2, E = 3, N = 4 };
enum { SW = 1, SE = 2, NE = 3, NW = 4 };

#define CN(d, c) cell_next[((c)-1) * 4 + ((d)-1)]
#define VOC(v, c) voc[((c)-1) * 4 + ((v)-1)]

/* panel_edge_cells(i, dir, panel)
 * (mesh_tools/source/support/gencube_ps_mod.F90:1943-1973) */
static int pec(int ec, int i, int dir, int panel) {
  int cpp = ec * ec;
  switch (dir) {
  case W: return (panel - 1) * cpp + ec * (i - 1) + 1;
  case S: return (panel - 1) * cpp + (cpp - ec) + i;
  case E: return (panel - 1) * cpp + ec * i;
  default: return (panel - 1) * cpp + i;
  }
}

/* calc_adjacency, calc_face_to_vert and orient_lfric
 * (mesh_tools/source/support/gencube_ps_mod.F90:336-546, 558-720, 1561-1601).
 * cell_next[c*4 + d] (W,S,E,N) and verts_on_cell[c*4 + v] (SW,SE,NE,NW) are 1-based ids. */
static int topology_impl(int ec, int *cell_next, int *verts_on_cell, int orient);
int orc_cubedsphere_topology(int ec, int *cell_next, int *verts_on_cell) {
  return topology_impl(ec, cell_next, verts_on_cell, 1);
}

static int topology_impl(int ec, int *cell_next, int *verts_on_cell, int orient) {
  if (ec < 1) return -1;
  const int cpp = ec * ec, ncells = 6 * cpp;
  int *voc = verts_on_cell;

  /* --- calc_adjacency (:336-546) --- */
  for (int c = 1; c <= ncells; ++c) {
    CN(W, c) = c - 1;
    CN(S, c) = c + ec;
    CN(E, c) = c + 1;
    CN(N, c) = c - ec;
  }
  /* panel_next(W,S,E,N, panel) (:390-395) */
  static const int pnext[7][5] = {{0},           {0, 4, 6, 2, 5}, {0, 1, 6, 3, 5},
                                  {0, 2, 6, 4, 5}, {0, 3, 6, 1, 5}, {0, 1, 2, 3, 4},
                                  {0, 1, 4, 3, 2}};
  /* per panel and side: which side of the neighbour panel, and whether the index runs
   * reversed (ec+1-i) — the six "Panel I..VI" loops at :408-542 */
  static const int side[7][5] = {{0},           {0, E, W, W, W}, {0, E, N, W, S},
                                 {0, E, E, W, E}, {0, E, S, W, N}, {0, N, N, N, N},
                                 {0, S, S, S, S}};
  static const int rev[7][5] = {{0},           {0, 0, 1, 0, 0}, {0, 0, 0, 0, 0},
                                {0, 0, 0, 0, 1}, {0, 0, 1, 0, 1}, {0, 0, 0, 1, 1},
                                {0, 1, 1, 0, 0}};
  for (int p = 1; p <= 6; ++p) {
    for (int i = 1; i <= ec; ++i) {
      for (int d = W; d <= N; ++d) {
        int cell = pec(ec, i, d, p);
        int j = rev[p][d] ? ec + 1 - i : i;
        CN(d, cell) = pec(ec, j, side[p][d], pnext[p][d]);
      }
    }
  }

  /* --- calc_face_to_vert (:558-720) --- */
  memset(voc, 0, sizeof(int) * 4 * (size_t)ncells);
  int nxf;
  for (int c = 1; c <= 4 * cpp; ++c) VOC(NW, c) = c;
  nxf = 4 * cpp + 1;
  for (int c = 1; c <= 4 * cpp; ++c) VOC(NE, c) = VOC(NW, CN(E, c));
  for (int p = 1; p <= 4; ++p)
    for (int c = (p - 1) * cpp + 1; c <= p * cpp - ec; ++c) {
      VOC(SW, c) = VOC(NW, CN(S, c));
      VOC(SE, c) = VOC(NE, CN(S, c));
    }
  for (int idx = 0; idx <= ec - 2; ++idx)
    for (int c = 4 * cpp + idx * ec + 1; c <= 4 * cpp + idx * ec + ec - 1; ++c) {
      VOC(SE, c) = nxf++;
      VOC(SW, CN(E, c)) = VOC(SE, c);
      VOC(NE, CN(S, c)) = VOC(SE, c);
      VOC(NW, CN(E, CN(S, c))) = VOC(SE, c);
    }
  for (int c = 5 * cpp + 1; c <= 6 * cpp; ++c) VOC(NW, c) = nxf++;
  for (int c = 5 * cpp + 1; c <= 6 * cpp - ec; ++c) VOC(SW, c) = VOC(NW, CN(S, c));
  for (int c = 6 * cpp - ec + 1; c <= 6 * cpp; ++c) VOC(SW, c) = nxf++;
  for (int c = 3 * cpp - ec + 1; c <= 3 * cpp; ++c) {
    VOC(SW, c) = nxf;
    VOC(SE, CN(W, c)) = VOC(SW, c);
    nxf++;
  }
  {
    int c = 3 * cpp;
    VOC(SE, c) = nxf;
    VOC(SW, CN(E, c)) = VOC(SE, c);
  }
  for (int c = cpp - ec + 1; c <= cpp; ++c) { /* S=>W join, I=>VI */
    VOC(SW, c) = VOC(SW, CN(S, c));
    VOC(SE, c) = VOC(NW, CN(S, c));
  }
  for (int c = 1; c <= ec; ++c) { /* N=>W join, I=>V */
    VOC(NW, CN(N, c)) = VOC(NW, c);
    VOC(SW, CN(N, c)) = VOC(NE, c);
  }
  for (int c = 2 * cpp + 1; c <= 2 * cpp + ec; ++c) { /* N=>E join, III=>V */
    VOC(SE, CN(N, c)) = VOC(NW, c);
    VOC(NE, CN(N, c)) = VOC(NE, c);
  }
  for (int c = 3 * cpp - ec + 1; c <= 3 * cpp; ++c) { /* E=>S join, III=>VI */
    VOC(NE, CN(S, c)) = VOC(SW, c);
    VOC(SE, CN(S, c)) = VOC(SE, c);
  }
  for (int c = 3 * cpp + 1; c <= 3 * cpp + ec; ++c) { /* N=>N join, IV=>V */
    VOC(NE, CN(N, c)) = VOC(NW, c);
    VOC(NW, CN(N, c)) = VOC(NE, c);
  }
  for (int idx = 0; idx <= ec - 1; ++idx)
    for (int c = 5 * cpp + 1 + idx * ec; c <= 5 * cpp + 1 + idx * ec + ec - 2; ++c) {
      VOC(NE, c) = VOC(NW, CN(E, c));
      VOC(SE, c) = VOC(SW, CN(E, c));
    }
  for (int c = 6 * cpp - ec + 1; c <= 6 * cpp; ++c) { /* S=>S join, VI=>IV */
    VOC(SW, CN(S, c)) = VOC(SE, c);
    VOC(SE, CN(S, c)) = VOC(SW, c);
  }
  for (int c = 2 * cpp - ec + 1; c <= 2 * cpp; ++c) { /* S=>N join, II=>VI */
    VOC(SW, c) = VOC(NW, CN(S, c));
    VOC(SE, c) = VOC(NE, CN(S, c));
  }
  for (int c = cpp + 1; c <= cpp + ec; ++c) { /* N=>S join, II=>V */
    VOC(SW, CN(N, c)) = VOC(NW, c);
    VOC(SE, CN(N, c)) = VOC(NE, c);
  }

  /* --- orient_lfric with PANEL_ROTATIONS = (0,0,1,1,-1,0) (:69, :1561-1601):
   * cshift(.,+1,dim=1) -> [a2,a3,a4,a1]; cshift(.,-1,dim=1) -> [a4,a1,a2,a3] --- */
  static const int rot[7] = {0, 0, 0, 1, 1, -1, 0};
  for (int p = 1; p <= 6 && orient; ++p) {
    if (rot[p] == 0) continue;
    for (int c = (p - 1) * cpp + 1; c <= p * cpp; ++c) {
      int *arrs[2] = {&voc[(c - 1) * 4], &cell_next[(c - 1) * 4]};
      for (int a = 0; a < 2; ++a) {
        int t[4];
        for (int i = 0; i < 4; ++i) t[i] = arrs[a][i];
        for (int i = 0; i < 4; ++i)
          arrs[a][i] = rot[p] == 1 ? t[(i + 1) % 4] : t[(i + 3) % 4];
      }
    }
  }
  return 0;
}

/* Local vertex ids: first encounter over cells, then the cell's vertices in stored order
 * (infrastructure/source/mesh/local_mesh_mod.f90:381-406).  Serial partition: local cell
 * id == global cell id (infrastructure/source/mesh/partition_mod.F90:620-628).
 * Returns the number of unique vertices. */
int orc_local_vert_lids(int ncells, const int *gid, int *lid) {
  int maxg = 0;
  for (int i = 0; i < 4 * ncells; ++i)
    if (gid[i] > maxg) maxg = gid[i];
  int *seen = calloc((size_t)maxg + 1, sizeof(int));
  if (!seen) return -1;
  int nu = 0;
  for (int i = 0; i < 4 * ncells; ++i) {
    if (!seen[gid[i]]) seen[gid[i]] = ++nu;
    lid[i] = seen[gid[i]];
  }
  free(seen);
  return nu;
}

/* mesh_connectivity: face and edge ids on each 3D cell of the base layer
 * (infrastructure/source/mesh/mesh_constructor_helper_functions_mod.F90:266-438).
 * face_on_cell[c*6+f] (W,S,E,N,B,T), edge_on_cell[c*12+e]
 * (WB,SB,EB,NB, SW,SE,NE,NW, WT,ST,ET,NT); vert_on_cell holds the 4 bottom vertex lids. */
int orc_mesh_connectivity(int nc, const int *cell_next, const int *voc, int *foc, int *eoc,
                          int *nfaces_2d, int *nedges_2d) {
  const int nfh = 4, neh = 4, nvh = 4, nf = 6, ne = 12;
  memset(foc, 0, sizeof(int) * nf * (size_t)nc);
  memset(eoc, 0, sizeof(int) * ne * (size_t)nc);
#define FOC(f, c) foc[((c)-1) * nf + (f)-1]
#define EOC(e, c) eoc[((c)-1) * ne + (e)-1]
#define NXT(f, c) cell_next[((c)-1) * 4 + (f)-1]
#define V2(v, c) voc[((c)-1) * 4 + (v)-1]
  int face_id = 1, nside = 0;
  for (int cell = 1; cell <= nc; ++cell) {
    for (int face = 1; face <= nfh; ++face) {
      if (FOC(face, cell) == 0) {
        FOC(face, cell) = face_id;
        int nbr = NXT(face, cell);
        if (nbr != 0)
          for (int fn = 1; fn <= nfh; ++fn)
            if (NXT(fn, nbr) == cell) FOC(fn, nbr) = face_id;
        face_id++;
        nside++;
      }
    }
    for (int face = nfh + 1; face <= nf; ++face) FOC(face, cell) = face_id++;
  }
  int edge_id = 1, nhedge = 0;
  for (int cell = 1; cell <= nc; ++cell) {
    for (int edge = 1; edge <= neh; ++edge) {
      if (EOC(edge, cell) == 0) {
        EOC(edge, cell) = edge_id;
        EOC(edge + neh + nvh, cell) = edge_id + 1;
        int nbr = NXT(edge, cell);
        if (nbr != 0)
          for (int en = 1; en <= neh; ++en)
            if (NXT(en, nbr) == cell) {
              EOC(en, nbr) = edge_id;
              EOC(en + neh + nvh, nbr) = edge_id + 1;
            }
        edge_id += 2;
        nhedge++;
      }
    }
    for (int edge = neh + 1; edge <= neh + nvh; ++edge) {
      if (EOC(edge, cell) == 0) {
        EOC(edge, cell) = edge_id;
        int vert = V2(edge - neh, cell);
        for (int face = 1; face <= nfh; ++face) {
          int nbr = NXT(face, cell);
          if (nbr == 0) continue;
          for (int vn = 1; vn <= nvh; ++vn) {
            if (V2(vn, nbr) != vert) continue;
            EOC(vn + neh, nbr) = edge_id;
            for (int fn = 1; fn <= nfh; ++fn) {
              int nn = NXT(fn, nbr);
              if (nn == 0) continue;
              for (int vnn = 1; vnn <= nvh; ++vnn)
                if (V2(vnn, nn) == vert) EOC(vnn + neh, nn) = edge_id;
            }
          }
        }
        edge_id++;
      }
    }
  }
  if (nfaces_2d) *nfaces_2d = nside;
  if (nedges_2d) *nedges_2d = nhedge;
  return 0;
#undef FOC
#undef EOC
#undef NXT
#undef V2
}

/* entity dof counts (infrastructure/source/function_space/
 * function_space_constructor_helper_functions_mod.F90:258-266 W0, :268-282 W1, :284-295 W2,
 * :391-400 W3; element_order_h = element_order_v = k) */
typedef struct {
  int vert, edge_h, edge_v, face_h, face_v, vol, cell;
} ndofs_t;

static int ndof_setup(int space, int k, ndofs_t *d) {
  memset(d, 0, sizeof(*d));
  switch (space) {
  case ORC_W0:
    d->vert = 1; d->edge_h = k; d->edge_v = k; d->face_h = k * k; d->face_v = k * k;
    d->vol = k * k * k; d->cell = (k + 2) * (k + 2) * (k + 2);
    return 0;
  case ORC_W1:
    d->edge_h = k + 1; d->edge_v = k + 1;
    d->face_h = (k + 1) * k + k * (k + 1); d->face_v = 2 * (k + 1) * k;
    d->vol = 2 * k * (k + 1) * k + k * k * (k + 1);
    d->cell = 2 * (k + 1) * (k + 2) * (k + 2) + (k + 2) * (k + 2) * (k + 1);
    return 0;
  case ORC_W2:
    d->face_h = (k + 1) * (k + 1); d->face_v = (k + 1) * (k + 1);
    d->vol = 2 * k * (k + 1) * (k + 1) + (k + 1) * (k + 1) * k;
    d->cell = 2 * (k + 2) * (k + 1) * (k + 1) + (k + 1) * (k + 1) * (k + 2);
    return 0;
  case ORC_W3:
    d->vol = (k + 1) * (k + 1) * (k + 1); d->cell = d->vol;
    return 0;
  }
  return -1;
}

int orc_ndf(int space, int order) {
  ndofs_t d;
  return ndof_setup(space, order, &d) ? -1 : d.cell;
}

/* dofmap_setup for a single (serial) partition, ndata = 1, nlayers-first
 * (infrastructure/source/function_space/function_space_constructor_helper_functions_mod.F90:
 * 1715-2300; owned branches only since every entity is owned in serial).
 * dofmap[c*ndf + i] is the 1-based dof at the bottom of the column. */
int orc_dofmap(int space, int order, int nc, int nl, int nv2, int ne2, const int *voc,
               const int *foc, const int *eoc, int *dofmap, int *undf) {
  ndofs_t d;
  if (ndof_setup(space, order, &d)) return -1;
  const int nhf = 4, nf = 6, ne = 12, nhe = 4, nv2d = 4, nv = 8;
  const int nvert_layer = 2 * nv2, nedge_layer = 2 * ne2 + nv2, nface_layer = ne2 + 2 * nc;
  int s0 = d.vert > 1 ? d.vert : 1, s1 = d.edge_h > d.edge_v ? d.edge_h : d.edge_v;
  int s2 = d.face_h > d.face_v ? d.face_h : d.face_v, s3 = d.vol;
  if (s1 < 1) s1 = 1;
  if (s2 < 1) s2 = 1;
  if (s3 < 1) s3 = 1;
  int *d0 = calloc((size_t)s0 * nvert_layer, sizeof(int));
  int *d1 = calloc((size_t)s1 * nedge_layer, sizeof(int));
  int *d2 = calloc((size_t)s2 * nface_layer, sizeof(int));
  int *d3 = calloc((size_t)s3 * nc, sizeof(int));
  if (!d0 || !d1 || !d2 || !d3) return -2;
#define D0(i, e) d0[((e)-1) * s0 + (i)-1]
#define D1(i, e) d1[((e)-1) * s1 + (i)-1]
#define D2(i, e) d2[((e)-1) * s2 + (i)-1]
#define D3(i, e) d3[((e)-1) * s3 + (i)-1]
#define FID(f, c) foc[((c)-1) * nf + (f)-1]
#define EID(e, c) eoc[((c)-1) * ne + (e)-1]
  /* 3D vertex ids: bottom = 2D lid, top = lid + nverts_2d (mesh_constructor_helper_functions_mod.F90:219-225) */
#define VID(v, c) ((v) <= nv2d ? voc[((c)-1) * 4 + (v)-1] : voc[((c)-1) * 4 + (v)-nv2d - 1] + nv2)
  const int off = 1; /* ndata_offset for nlayers-first data (:1833-1837) */
  int id = 1;
  for (int c = 1; c <= nc; ++c) {
    for (int i = 1; i <= d.vol; ++i) { /* (3,3) */
      D3(i, c) = id;
      id += nl;
    }
    for (int f = 1; f <= nhf; ++f) { /* side faces */
      int fid = FID(f, c);
      if (D2(1, fid) == 0)
        for (int i = 1; i <= d.face_h; ++i) {
          D2(i, fid) = id;
          id += nl;
        }
    }
    int id0 = id; /* bottom / top faces */
    for (int f = nhf + 1; f <= nf; ++f) {
      int fid = FID(f, c);
      if (D2(1, fid) == 0)
        for (int i = 1; i <= d.face_v; ++i) {
          D2(i, fid) = id;
          id += nl + 1;
        }
      if (f == nhf + 1) id = id0 + off;
      else id = id - off;
    }
    for (int e = 1; e <= nhe; ++e) { /* horizontal edges */
      int be = EID(e, c), te = EID(e + ne - nhe, c);
      if (D1(1, be) == 0)
        for (int i = 1; i <= d.edge_h; ++i) {
          D1(i, be) = id;
          D1(i, te) = id + off;
          id += nl + 1;
        }
    }
    for (int e = nhe + 1; e <= ne - nhe; ++e) { /* vertical edges */
      int se = EID(e, c);
      if (D1(1, se) == 0)
        for (int i = 1; i <= d.edge_v; ++i) {
          D1(i, se) = id;
          id += nl;
        }
    }
    for (int v = 1; v <= nv2d; ++v) { /* vertices */
      int bv = VID(v, c), tv = VID(v + nv2d, c);
      if (D0(1, bv) == 0)
        for (int i = 1; i <= d.vert; ++i) {
          D0(i, bv) = id;
          D0(i, tv) = id + off;
          id += nl + 1;
        }
    }
  }
  *undf = id - 1;
  /* copy into one cell dofmap in entity order vol, faces, edges, vertices (:2182-2283) */
  for (int c = 1; c <= nc; ++c) {
    int *row = &dofmap[(size_t)(c - 1) * d.cell];
    int k = 0;
    for (int i = 1; i <= d.vol; ++i)
      if (D3(i, c) != 0) row[k++] = D3(i, c);
    for (int f = 1; f <= nf; ++f) {
      int stop = f <= nhf ? d.face_h : d.face_v;
      for (int i = 1; i <= stop; ++i)
        if (D2(i, FID(f, c)) != 0) row[k++] = D2(i, FID(f, c));
    }
    for (int e = 1; e <= ne; ++e) {
      int stop = (e <= nhe || e > ne - nhe) ? d.edge_h : d.edge_v;
      for (int i = 1; i <= stop; ++i)
        if (D1(i, EID(e, c)) != 0) row[k++] = D1(i, EID(e, c));
    }
    for (int v = 1; v <= nv; ++v)
      for (int i = 1; i <= d.vert; ++i)
        if (D0(i, VID(v, c)) != 0) row[k++] = D0(i, VID(v, c));
    if (k != d.cell) {
      free(d0); free(d1); free(d2); free(d3);
      return -3;
    }
  }
  free(d0); free(d1); free(d2); free(d3);
  return 0;
#undef D0
#undef D1
#undef D2
#undef D3
#undef FID
#undef EID
#undef VID
}

/* Closed-form 4-colouring of a cubed sphere (infrastructure/source/mesh/mesh_colouring_mod.F90:
 * 151-308, tables :242-247 even n and :263-268 odd n) and its validator (:765-837).
 * Fills colour_map; returns 0, or 1 when the closed form does not validate (the reference then
 * falls back to set_colours_generic, :100-113). */
int orc_colour_cubedsphere(int n, const int *cell_next, int *colour_map) {
  const int cpp = n * n, nc = 6 * cpp;
  /* cell_colour_order_per_panel(column,row,panel), reshape((/a,b,c,d/),(/2,2/)) ->
   * (1,1)=a (2,1)=b (1,2)=c (2,2)=d */
  static const int even[7][4] = {{0},        {1, 2, 3, 4}, {3, 2, 1, 4}, {3, 4, 1, 2},
                                 {1, 4, 3, 2}, {3, 2, 4, 1}, {2, 3, 1, 4}};
  static const int odd[7][4] = {{0},        {1, 2, 3, 4}, {4, 1, 2, 3}, {1, 2, 3, 4},
                                {4, 1, 2, 3}, {3, 2, 4, 1}, {3, 2, 4, 1}};
  const int(*tab)[4] = (n % 2 == 0) ? even : odd;
  for (int cell = 1; cell <= nc; ++cell) {
    int panel = (cell - 1) / cpp + 1;
    int rel = cell - cpp * (panel - 1);
    int column = 2 - (rel % n) % 2;
    int row = ((rel - 1) / n) % 2 + 1;
    colour_map[cell - 1] = tab[panel][(column - 1) + 2 * (row - 1)];
  }
  return orc_colour_validate(nc, cell_next, colour_map) ? 0 : 1;
}

/* set_colours_generic (mesh_colouring_mod.F90:514-650) with choose_colour_greedy (:848-864):
 * cells in id order take the lowest colour (of at most 50) not yet held by a cell that shares
 * an edge or, through the N and S neighbours, a vertex.  Fills colour_map; returns the number
 * of colours (the first colour with no cells, minus one), or 0 when the result does not
 * validate. */
int orc_colour_generic(int nc, const int *cell_next, int *colour_map) {
  enum { MAXCOLS = 50 };
  int count[MAXCOLS + 1] = {0};
#define NX(d, c) cell_next[((c)-1) * 4 + (d)-1]
#define CM(c) ((c) > 0 ? colour_map[(c)-1] : 0)
  for (int cell = 1; cell <= nc; ++cell) colour_map[cell - 1] = 0;
  for (int cell = 1; cell <= nc; ++cell) {
    int used[MAXCOLS + 1] = {0};
    int t = NX(N, cell);
    if (t > 0) {
      used[CM(t)] = 1;
      if (cell == NX(S, t)) {
        used[CM(NX(E, t))] = 1;
        used[CM(NX(W, t))] = 1;
      } else {
        used[CM(NX(N, t))] = 1;
        used[CM(NX(S, t))] = 1;
      }
    }
    t = NX(E, cell);
    if (t > 0) used[CM(t)] = 1;
    t = NX(W, cell);
    if (t > 0) used[CM(t)] = 1;
    t = NX(S, cell);
    if (t > 0) {
      used[CM(t)] = 1;
      if (cell == NX(N, t)) {
        used[CM(NX(E, t))] = 1;
        used[CM(NX(W, t))] = 1;
      } else {
        used[CM(NX(N, t))] = 1;
        used[CM(NX(S, t))] = 1;
      }
    }
    int colour = 0;
    for (int i = 1; i <= MAXCOLS; ++i)
      if (!used[i]) {
        colour = i;
        break;
      }
    colour_map[cell - 1] = colour;
    count[colour]++;
  }
#undef NX
#undef CM
  int ncolours = MAXCOLS;
  for (int i = 1; i <= MAXCOLS; ++i)
    if (count[i] == 0) {
      ncolours = i - 1;
      break;
    }
  return orc_colour_validate(nc, cell_next, colour_map) ? ncolours : 0;
}

/* create_colour_map_lookup (mesh_colouring_mod.F90:703-752): cells of each colour in id
 * order.  cic is [ncolours][maxcc] row-major by colour, zero-padded; maxcc must be at least
 * the largest colour count.  Returns that count. */
int orc_colour_lookup(int nc, int ncolours, const int *colour_map, int *npc, int *cic, int maxcc) {
  int mx = 0;
  for (int c = 0; c < ncolours; ++c) npc[c] = 0;
  for (int cell = 1; cell <= nc; ++cell) npc[colour_map[cell - 1] - 1]++;
  for (int c = 0; c < ncolours; ++c) mx = npc[c] > mx ? npc[c] : mx;
  if (mx > maxcc) return mx;
  for (int i = 0; i < ncolours * maxcc; ++i) cic[i] = 0;
  for (int c = 0; c < ncolours; ++c) {
    int i = 0;
    for (int cell = 1; cell <= nc; ++cell)
      if (colour_map[cell - 1] == c + 1) cic[c * maxcc + i++] = cell;
  }
  return mx;
}

/* test_mesh_planar_quadrilateral and set_colours_planar_quadrilateral
 * (infrastructure/source/mesh/mesh_colouring_mod.F90:663-691, 348-495): when every neighbour
 * leads back across the opposite side, a breadth-first walk from cell 1 (then from the next
 * uncoloured cell) over the N, E, S, W neighbours of each front, each neighbour taking
 * cell_next_colour(direction, colour); 4 colours.  Returns 4, or 0 when the mesh is not planar
 * quadrilateral or the result does not validate (the reference then colours generically). */
int orc_colour_planar(int nc, const int *cell_next, int *colour_map) {
  static const int cell_next_colour[5][5] = {{0, 0, 0, 0, 0}, {0, 2, 1, 4, 3}, {0, 3, 4, 1, 2},
                                             {0, 2, 1, 4, 3}, {0, 3, 4, 1, 2}};  /* [dir W..N][colour] */
  static const int direction[4] = {N, E, S, W}, opposite[4] = {S, W, N, E};
#define NX(d, c) cell_next[((c)-1) * 4 + (d)-1]
  for (int cell = 1; cell <= nc; ++cell)
    for (int k = 0; k < 4; ++k) {
      const int t = NX(direction[k], cell);
      if (t > 0 && t <= nc && NX(opposite[k], t) != cell) return 0;
    }
  int *todo = (int *)calloc((size_t)nc + 1, sizeof(int));
  for (int c = 0; c < nc; ++c) colour_map[c] = 0;
  int cell = 1, start = 1, end = 1;
  todo[1] = 1;
  while (cell <= nc) {
    colour_map[cell - 1] = 1;
    while (start <= end) {
      int i;
      const int last = end;
      for (i = start; i <= last; ++i) {
        const int c = todo[i];
        for (int k = 0; k < 4; ++k) {
          const int t = NX(direction[k], c);
          if (t > 0 && t <= nc && colour_map[t - 1] == 0) {
            colour_map[t - 1] = cell_next_colour[direction[k]][colour_map[c - 1]];
            todo[++end] = t;
          }
        }
      }
      start = i;
    }
    while (colour_map[cell - 1] != 0) {
      ++cell;
      if (cell > nc) break;
    }
  }
  free(todo);
#undef NX
  return orc_colour_validate(nc, cell_next, colour_map) ? 4 : 0;
}

/* test_set_colours (infrastructure/source/mesh/mesh_colouring_mod.F90:765-837).
 * Returns 1 when the colouring passes. */
int orc_colour_validate(int nc, const int *cell_next, const int *cm) {
#define NX(d, c) cell_next[((c)-1) * 4 + (d)-1]
#define CM(c) ((c) > 0 ? cm[(c)-1] : 0)
  for (int cell = 1; cell <= nc; ++cell) {
    int col = cm[cell - 1];
    if (col == CM(NX(N, cell)) || col == CM(NX(S, cell)) || col == CM(NX(E, cell)) ||
        col == CM(NX(W, cell)))
      return 0;
    int t = NX(N, cell);
    if (t > 0) {
      if (cell == NX(S, t)) {
        if (col == CM(NX(E, t)) || col == CM(NX(W, t))) return 0;
      } else if (col == CM(NX(N, t)) || col == CM(NX(S, t))) {
        return 0;
      }
    }
    t = NX(S, cell);
    if (t > 0) {
      if (cell == NX(N, t)) {
        if (col == CM(NX(E, t)) || col == CM(NX(W, t))) return 0;
      } else if (col == CM(NX(N, t)) || col == CM(NX(S, t))) {
        return 0;
      }
    }
  }
  return 1;
#undef NX
#undef CM
}

/* ---- coordinates ------------------------------------------------------------------------ */
#include <float.h>
#include <math.h>

/* xyz2ll_r_double (infrastructure/source/utilities/coord_transform_mod.F90:226-283) */
static void xyz2ll(double x, double y, double z, double *longitude, double *latitude) {
  const double pi = 4.0 * atan(1.0);
  if (x == 0.0) {
    *longitude = y >= 0.0 ? 0.5f * pi : -0.5f * pi;
  } else {
    *longitude = atan(y / x);
    if (x < 0.0) {
      if (y > 0.0)
        *longitude = *longitude + pi;
      else
        *longitude = *longitude - pi;
    }
  }
  const double radius = sqrt(x * x + y * y);
  if (radius <= DBL_EPSILON) {
    *latitude = z > 0.0 ? 0.5 * pi : -0.5 * pi;
    *longitude = 0.0;
  } else {
    *latitude = atan(z / radius);
  }
}

/* calc_coords (mesh_tools/source/support/gencube_ps_mod.F90:887-1113), called by generate
 * (:1302-1348) after calc_face_to_vert and before orient_lfric: vert_coords(1,v) = longitude,
 * vert_coords(2,v) = -latitude (radians), v the global vertex id (1-based; lon[v-1]). */
int orc_cubedsphere_lonlat(int ec, double *lon, double *lat) {
  if (ec < 1) return -1;
  const int cpp = ec * ec, ncells = 6 * cpp;
  int *cn = malloc(sizeof(int) * 4 * (size_t)ncells), *voc = malloc(sizeof(int) * 4 * (size_t)ncells);
  if (!cn || !voc) {
    free(cn);
    free(voc);
    return -1;
  }
  topology_impl(ec, cn, voc, 0);
  const double pi = 4.0 * atan(1.0), pio4 = pi / 4.0;
  const double dlambda = 0.5 * pi / ec;
  double lo, la, xs, ys, zs, x0, y0, z0, t1, t2, lambda1, lambda2;
#define SET(vid, X, Y, Z)            \
  do {                               \
    xyz2ll((X), (Y), (Z), &lo, &la); \
    lon[(vid)-1] = lo;               \
    lat[(vid)-1] = -la;              \
  } while (0)
#define CUBE()                                        \
  do {                                                \
    xs = 1.0 / sqrt(1.0 + t1 * t1 + t2 * t2);         \
    ys = xs * t1;                                     \
    zs = xs * t2;                                     \
  } while (0)
  int vert = 1;
  for (int y = 1; y <= ec; ++y) { /* Panels I-IV */
    lambda2 = (y - 1) * dlambda - pio4;
    t2 = tan(lambda2);
    for (int x = 1; x <= ec; ++x) {
      lambda1 = (x - 1) * dlambda - pio4;
      t1 = tan(lambda1);
      CUBE();
      SET(vert, xs, ys, zs);
      x0 = -ys; y0 = xs; z0 = zs;
      SET(vert + cpp, x0, y0, z0);
      x0 = -xs; y0 = -ys; z0 = zs;
      SET(vert + 2 * cpp, x0, y0, z0);
      x0 = ys; y0 = -xs; /* z0 keeps Panel III's value */
      SET(vert + 3 * cpp, x0, y0, z0);
      vert++;
    }
  }
  int cell = 4 * cpp; /* Panel V */
  for (int y = 1; y <= ec - 1; ++y) {
    lambda2 = y * dlambda - pio4;
    t2 = tan(lambda2);
    for (int x = 1; x <= ec - 1; ++x) {
      lambda1 = x * dlambda - pio4;
      t1 = tan(lambda1);
      CUBE();
      SET(VOC(SE, cell + x), -ys, zs, -xs);
    }
    cell += ec;
  }
  cell = 5 * cpp + 1; /* Panel VI */
  for (int y = 1; y <= ec; ++y) {
    lambda2 = (y - 1) * dlambda - pio4;
    t2 = tan(lambda2);
    for (int x = 1; x <= ec; ++x) {
      lambda1 = (x - 1) * dlambda - pio4;
      t1 = tan(lambda1);
      CUBE();
      SET(VOC(NW, cell), -ys, -zs, xs);
      cell++;
    }
  }
  cell = 6 * cpp - ec + 1; /* Panel VI: bottom edge */
  lambda2 = ec * dlambda - pio4;
  t2 = tan(lambda2);
  for (int x = 1; x <= ec; ++x) {
    lambda1 = (x - 1) * dlambda - pio4;
    t1 = tan(lambda1);
    CUBE();
    SET(VOC(SW, cell), -ys, -zs, xs);
    cell++;
  }
  cell = 5 * cpp + ec; /* Panel VI: right edge */
  lambda1 = ec * dlambda - pio4;
  t1 = tan(lambda1);
  for (int y = 1; y <= ec; ++y) {
    lambda2 = (y - 1) * dlambda - pio4;
    t2 = tan(lambda2);
    CUBE();
    SET(VOC(NE, cell), -ys, -zs, xs);
    cell += ec;
  }
  lambda1 = ec * dlambda - pio4;
  t1 = tan(lambda1);
  lambda2 = ec * dlambda - pio4;
  t2 = tan(lambda2);
  CUBE();
  SET(6 * cpp + 2, -ys, -zs, xs);
#undef SET
#undef CUBE
  free(cn);
  free(voc);
  return 0;
}

/* uniform_extrude / quadratic_extrude / geometric_extrude with geometric_func
 * (infrastructure/source/mesh/extrusion_mod.f90:292-305, 341-354, 390-409, 618-629); kind 0 uniform,
 * 1 quadratic, 2 geometric.  eta has nl+1 entries. */
int orc_eta(int kind, int nl, double *eta) {
  if (nl < 1) return -1;
  if (kind == 0) {
    for (int k = 0; k <= nl; ++k) eta[k] = (double)k / (double)nl;
  } else if (kind == 1) {
    for (int k = 0; k <= nl; ++k) {
      const double q = (double)k / (double)nl;
      eta[k] = q * q;
    }
  } else if (kind == 2) {
    const double sf = 1.03;
    double p = 1.0;
    for (int i = 0; i < nl; ++i) p = p * sf; /* stretching_factor**number_of_layers */
    const double delta_eta = (sf - 1.0) / (p - 1.0);
    eta[0] = 0.0;
    for (int k = 1; k <= nl; ++k) eta[k] = delta_eta * (pow(sf, (double)k) - 1.0) / (sf - 1.0);
  } else {
    return -1;
  }
  return 0;
}

/* Cartesian coordinate fields on lowest-order W0 (dofmap(8, ncells), 1-based columns), as
 * assign_coordinate_field's xyz branch (components/driver/source/driver_coordinates_mod.F90:
 * 357-454) leaves them for a mesh extruded by mesh_extruder (infrastructure/source/mesh/
 * mesh_constructor_helper_functions_mod.F90:120-252) with dz from set_dz (:450-477) over a zero
 * surface height: the cells are visited in order, layers k = 0..nl-1 and the 8 nodes in order,
 * and every node is (re)written, so the last write stands.  vlon/vlat by the vertex id in
 * vert_on_cell(4, ncells). */
int orc_chi_w0(int ncells, int nl, const int *dofmap, const int *vert_on_cell, const double *vlon,
               const double *vlat, int kind, double domain_top, double scaled_radius, double *chi1,
               double *chi2, double *chi3) {
  double *eta = malloc(sizeof(double) * (size_t)(nl + 1)), *dz = malloc(sizeof(double) * (size_t)nl);
  if (!eta || !dz || orc_eta(kind, nl, eta)) {
    free(eta);
    free(dz);
    return -1;
  }
  for (int k = 1; k <= nl; ++k) dz[k - 1] = (eta[k] - eta[k - 1]) * domain_top;
  static const double hat[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int c = 0; c < ncells; ++c) {
    for (int k = 0; k < nl; ++k) {
      double vc[8][3];
      for (int v = 0; v < 8; ++v) { /* column_coords: the cell's vertices of layer k */
        const int vid = vert_on_cell[c * 4 + (v % 4)] - 1;
        const int lev = k + v / 4;
        double s = 0.0;
        for (int j = 1; j <= lev; ++j) s = s + dz[j - 1];
        const double r = 1.0 + s;
        const double sl = sin(vlon[vid]), cl = cos(vlon[vid]), sa = sin(vlat[vid]), ca = cos(vlat[vid]);
        vc[v][0] = r * cl * ca;
        vc[v][1] = r * sl * ca;
        vc[v][2] = r * sa;
      }
      for (int df = 0; df < 8; ++df) {
        double x = 0.0, y = 0.0, z = 0.0;
        for (int v = 0; v < 8; ++v) {
          const double w = (1.0 - fabs(hat[v][0] - hat[df][0])) * (1.0 - fabs(hat[v][1] - hat[df][1])) *
                           (1.0 - fabs(hat[v][2] - hat[df][2]));
          x = x + w * vc[v][0];
          y = y + w * vc[v][1];
          z = z + w * vc[v][2];
        }
        double s = 0.0;
        for (int j = 1; j <= k; ++j) s = s + dz[j - 1];
        double rc = scaled_radius + s + hat[df][2] * dz[k];
        rc = rc / sqrt(x * x + y * y + z * z);
        const int dfk = dofmap[c * 8 + df] - 1 + k;
        chi1[dfk] = x * rc;
        chi2[dfk] = y * rc;
        chi3[dfk] = z * rc;
      }
    }
  }
  free(eta);
  free(dz);
  return 0;
}
