/* HU -> tissue classes and density on the GPU: the producer of the density map
 * that the dose epilogue (include/pvd.h) consumes.
 *
 * Replaces TissueComposition.calculate_composition and _handle_artifacts
 * (pyvoxeldosimetry/tissue/composition.py:48-94): an optional metal-artifact
 * fill (voxels above `metal_hu` take the value of a 3-D Gaussian smoothing of
 * the image in which metal and NaN voxels are 0, scipy.ndimage.gaussian_filter
 * semantics: 'reflect' boundaries, radius = (int)(truncate * sigma + 0.5)),
 * then inclusive HU-range class masks (composition.py:40-46, :63-68).  It also
 * maps HU to mass density through a piecewise-constant material table (the
 * ranges of core/gate/data/HU_to_material.txt:5-19 with densities of
 * core/gate/data/GateMaterials.db).
 *
 * Layout: C-order (n0, n1, n2) float32 HU, z contiguous.  All pointers are
 * device pointers; every output may be NULL (not produced).  Stream-ordered,
 * allocates nothing.  Returns PVD_OK or a negative PVD_E* code (include/pvd.h). */
#ifndef PVD_TISSUE_H
#define PVD_TISSUE_H

#include "pvd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PVD_TISSUE_MAX_CLASSES 8
#define PVD_TISSUE_MAX_MATERIALS 32
#define PVD_TISSUE_MAX_RADIUS 16

typedef struct {
    int n_classes;                                   /* <= PVD_TISSUE_MAX_CLASSES                    */
    float class_lo[PVD_TISSUE_MAX_CLASSES];          /* bit c of class_bits[v] is set when            */
    float class_hi[PVD_TISSUE_MAX_CLASSES];          /*   class_lo[c] <= hu <= class_hi[c]            */
    int n_materials;                                 /* <= PVD_TISSUE_MAX_MATERIALS, >= 1             */
    float mat_lo[PVD_TISSUE_MAX_MATERIALS];          /* ascending; material i covers [lo_i, lo_{i+1}),*/
    float mat_density[PVD_TISSUE_MAX_MATERIALS];     /*   material 0 also everything below lo_0 (g/cc) */
    int handle_artifacts;                            /* 1: metal fill before classification           */
    float metal_hu;                                  /* metal: hu > metal_hu (reference: 2000)         */
    double sigma;                                    /* Gaussian sigma in voxels (reference: 1)        */
    double truncate;                                 /* scipy's truncate (default 4)                   */
} pvd_tissue_params;

/* corrected[v]: HU after the metal fill (input HU where no fill applies);
 * class_bits[v]: class mask bits; density[v]: g/cm3 of the material of
 * corrected[v] (NaN for NaN HU). */
int pvd_tissue_map(const pvd_tissue_params *p, const float *ct, int n0, int n1, int n2, float *corrected,
                   unsigned char *class_bits, float *density, void *stream);

#ifdef __cplusplus
}
#endif
#endif
