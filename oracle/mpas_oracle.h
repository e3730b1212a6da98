/*
 * CPU oracle for the split-explicit dynamics step (atm_srk3).
 *
 * TEST INFRASTRUCTURE ONLY.  This restatement is the checker for the HIP
 * path and the timed CPU baseline; nothing in the product links it.
 *
 * Layout follows the reference runtime tensors (docs/API.md "State and
 * controls"): row-major (rows, levels) with one trailing padding row per
 * entity; connectivity here is 0-based with -1 for a missing neighbour
 * (the Python binding converts the reference's 1-based values).
 */
#ifndef MPAS_ORACLE_H
#define MPAS_ORACLE_H

#include <stdint.h>

typedef struct orc_mesh {
    int32_t nCells, nEdges, nVertices, nLevels;
    int32_t maxEdges, maxEdges2, vertexDegree, maxAdv;
    /* connectivity, 0-based, -1 = none */
    const int32_t *cellsOnEdge;      /* [E][2] */
    const int32_t *verticesOnEdge;   /* [E][2] */
    const int32_t *nEdgesOnCell;     /* [C] */
    const int32_t *edgesOnCell;      /* [C][maxEdges] */
    const int32_t *cellsOnCell;      /* [C][maxEdges] */
    const int32_t *verticesOnCell;   /* [C][maxEdges] */
    const int32_t *kiteForCell;      /* [C][maxEdges] 0-based slot */
    const int32_t *edgesOnVertex;    /* [V][3] */
    const int32_t *nEdgesOnEdge;     /* [E] */
    const int32_t *edgesOnEdge;      /* [E][maxEdges2] */
    const int32_t *nAdvCellsForEdge; /* [E] */
    const int32_t *advCellsForEdge;  /* [E][maxAdv] */
    /* horizontal metrics */
    const double *edgesOnCell_sign;   /* [C][maxEdges] */
    const double *edgesOnVertex_sign; /* [V][3] */
    const double *weightsOnEdge;      /* [E][maxEdges2] */
    const double *kiteAreasOnVertex;  /* [V][3] */
    const double *dcEdge, *dvEdge, *invDcEdge, *invDvEdge;
    const double *invAreaCell, *invAreaTriangle;
    const double *fEdge, *fVertex;
    const double *meshScalingDel2, *meshScalingDel4;
    const double *adv_coefs, *adv_coefs_3rd; /* [E][maxAdv] */
    const double *defc_a, *defc_b;           /* [C][maxEdges] */
    const double *coeffs_reconstruct;        /* [C][maxEdges][3] */
    const double *latCell, *lonCell;
    /* vertical metrics */
    const double *zz;      /* [C][L] */
    const double *zxu;     /* [E][L] */
    const double *dss;     /* [C][L] */
    const double *zb_cell, *zb3_cell; /* [C][maxEdges][L+1] */
    const double *rdzw, *rdzu, *fzm, *fzp, *rdnw, *fnm, *fnp; /* [L] */
    double cf1, cf2, cf3;
    /* base state */
    const double *rho_base, *rtheta_base, *exner_base; /* [C][L] */
} orc_mesh;

int orc_mesh_sizeof(void);

#endif
