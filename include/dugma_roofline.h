/* dugma_roofline.h -- canonical algorithmic FLOP per point pair of the pair
 * kernels, the numerator of every roofline figure reported for this library
 * (achieved FLOP/s = pairs per second x FLOP per pair).
 *
 * Counting: an FMA is 2 FLOP; each exp, rsq and rcp is 1.  The per-source
 * preparation (R x_j + t and R S_j R^T) is O(M) and not counted.
 *   3-D full covariance, forward (59): d 3, S = S_i + S_j 6, cofactors 18,
 *     det 5, adj(S) d 15, d^T adj(S) d 5, rcp 1, scale 2, exp 1, rsq 1,
 *     product 1, accumulate 1.
 *   3-D full covariance, adjoint (+48): w = S^-1 d 3, e w 3, S^-1 6,
 *     dE/dS 18, target-side accumulators 9, source-side accumulators 9.
 *   3-D isotropic: forward 17, adjoint +19.
 *   2-D full covariance: forward 24, adjoint +26.
 *   k-NN distance: 3 differences, 1 product, 2 FMA (plus compare / insert).
 *   compi-compatible pair pass: the weight P (two quadratic forms, two exps)
 *     about 40, plus 20 FMA feature accumulations.
 *   Batched registration pair (isotropic, sigma-update): S 2 (one FMA),
 *     d 3, |d|^2 5, rsq 1, exponent and weights 6 (with the exp),
 *     accumulations 7 (W and three FMA; the energy is summed in one pass
 *     after the iterations), nearest-distance min 1.  (The difference is
 *     counted as 3, as everywhere in this table: the FP32 hi/lo compensation
 *     some kernels add to it is not algorithmic work.)
 * The reference's own compi spends about 4.7 K operations per pair; it is
 * reported as a ratio, not against this table. */
#ifndef DUGMA_ROOFLINE_H
#define DUGMA_ROOFLINE_H

#define DUGMA_FLOP_3D_FULL_FWD 59
#define DUGMA_FLOP_3D_FULL_FUSED 107
#define DUGMA_FLOP_3D_ISO_FWD 17
#define DUGMA_FLOP_3D_ISO_FUSED 36
#define DUGMA_FLOP_2D_FULL_FWD 24
#define DUGMA_FLOP_2D_FULL_FUSED 50
#define DUGMA_FLOP_KNN 8
#define DUGMA_FLOP_COMPI 80
#define DUGMA_FLOP_BATCH_PAIR 25

/* FP32 vector peak of one MI355X, TFLOP/s (AMD's public figure:
 * 256 CU x 2.4 GHz x 256 FLOP per clock per CU). */
#define DUGMA_PEAK_FP32_TFLOPS 157.3

#endif /* DUGMA_ROOFLINE_H */
