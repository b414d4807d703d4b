/*
 * FP64 distance / ANI formulas evaluated on the four per-pair counts.
 * Compiled both for the device (HIP, the fused epilogue) and for the host
 * (plain C, the CLI's detail rows).  Evaluation order follows the reference
 * expressions so host and device agree to rounding.
 *
 *   naive distance         command_ani.h:200-224 (NUM_CODENS == 11 constants)
 *   p_dist                 command_ani.h:226-232
 *   CtxMoE (lm3ways)       model_ani.h:243-305 (coefficients model_ani.h:10-118)
 *   blastn-like AF model   model_ani.h:183-213, :307-367
 *   Mash / AAF distance    command_matrix.c:26-33 with command_matrix.h:10-13
 * The coefficient tables are the reference's fitted model parameters; they
 * have to be identical for the outputs to agree, so they are carried as data.
 */
#ifndef KSSD_MODELS_H
#define KSSD_MODELS_H
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define KSSD_HD __host__ __device__ static inline
#define KSSD_CONST static constexpr
#else
#define KSSD_HD static inline
#define KSSD_CONST static const
#endif

#define KSSD_EPS 1e-8

KSSD_CONST double kssd_p_main[12] = {
	1.0178887, 1.0135146, 1.0302181, 1.1205122,
	1.0136223, 0.9803396, 1.0021964, 0.9839395,
	0.9988935, 0.9690589, 0.9888171, 0.9817371};
KSSD_CONST double kssd_p_n95[12] = {
	1.0145656, 1.2708196, 1.2169829, 0.3904416,
	0.8147527, 1.3389150, 2.0581364, -1.4624736,
	1.6877678, 2.4390067, 0.7878482, 0.6697437};
KSSD_CONST double kssd_c_main[17] = {
	6.86905236142636, -5.85271800517096e-08, 7.43351941016881e-06,
	-1.95859228059721e-05, -8.44336369105416e-06, -11726.5478433625,
	-111109.682632437, 121000.895632267, -741.538955661909,
	-1894.53741972785, 2585.81964981999, -7744.5216892858,
	-18190.2514369096, 23306.3090989115, 8352.82760552325,
	19782.4294837312, -25518.2493259614};
KSSD_CONST double kssd_c_n95[17] = {
	-1.69173585242774, -3.64716315423252e-08, 2.85880264052975e-05,
	-5.5597924264054e-05, -2.96208613804271e-05, 1106.96550282022,
	-83.197177767225, -1671.88720051458, 152.75590068948,
	212.734805232395, 208.388739816682, -0.27860125939004,
	2.27065806912198, -10.2977013026837, -250.681324281127,
	-403.751360812353, -137.96628693833};
KSSD_CONST double kssd_p_af[15] = {
	2.36775520439829, 2.17036643548818, -0.0977124271437818, -0.251063182363729, 2.19491057829433,
	1.41850238263675, 2.15224961402778, -2.18587619877498, -2.21943802445092, 0.418474287083788,
	0.990424249133548, 0.502163745411501, 5.7787290524088, 0.626839640715682, 0.638706160147675};
KSSD_CONST double kssd_c_af[21] = {
	0.127498260534708, -1.51730137145957e-06, 9.48114616179122e-05,
	-0.000257916039708839, -7.85253629216324e-05, 1.96482915215685e-06,
	-1475.58996322769, 26.4698377745167, 529.387641078915,
	42.0602258298213, 92.3637001980455, -306.00755719743,
	-86.3482066935661, 2.16630645286942, -0.674746089642152,
	48.1147600903939, 3.39101087242059, -15.5481688297158,
	-26.7068442368566, 18.5142668221724, 20.4430437567379};

/* Count-like model inputs are scaled to the fold the models were fitted at
 * (DimRdcFold 8): 2^(drfold-8), clamped to 2^+-60. */
KSSD_HD double kssd_fold_scale(int drfold)
{
	int d = drfold - 8;
	if (d < -60)
		d = -60;
	if (d > 60)
		d = 60;
	return ldexp(1.0, d);
}

KSSD_HD double kssd_naive_dist(uint32_t xny, uint32_t dobj, uint32_t dsec)
{
	if (xny == 0)
		return 1;
	const double ratio = ((double)dsec + KSSD_EPS) / ((double)dobj + KSSD_EPS);
	const double dist0 = (double)dobj / (double)(uint32_t)(xny + dobj);
	const double fd = 1 - pow(1 - dist0, ratio);
	if (fd <= 0)
		return 0;
	double w = fd / 0.01;
	if (w > 1.0)
		w = 1.0;
	return fd * 0.1544286 + 7.133e-4 * w;
}

KSSD_HD double kssd_p_dist(uint32_t xny, uint32_t dsec, unsigned obj_bits)
{
	if (xny == 0 || obj_bits == 0)
		return 1;
	return (double)dsec / ((double)xny * ((double)obj_bits / 2.0));
}

KSSD_HD double kssd_lm_core(double x, double s, double m2, double d, int n95)
{
	const double *p = n95 ? kssd_p_n95 : kssd_p_main;
	const double *c = n95 ? kssd_c_n95 : kssd_c_main;
	const double d1 = 1 / (p[0] * x + p[1] * s + p[2] * m2 + p[3] * d + KSSD_EPS);
	const double d2 = 1 / (p[4] * x + p[5] * s + p[6] * m2 + p[7] * d + KSSD_EPS);
	const double d3 = 1 / (p[8] * x + p[9] * s + p[10] * m2 + p[11] * d + KSSD_EPS);
	return c[0] + c[1] * x + c[2] * s + c[3] * m2 + c[4] * d + c[5] * d1 + c[6] * d2 +
	       c[7] * d3 + c[8] * x * d1 + c[9] * s * d1 + c[10] * m2 * d1 + c[11] * x * d2 +
	       c[12] * s * d2 + c[13] * m2 * d2 + c[14] * x * d3 + c[15] * s * d3 + c[16] * m2 * d3;
}

KSSD_HD double kssd_moe_dist(uint32_t xny, uint32_t dobj, uint32_t dsec, uint32_t mut2, int drfold)
{
	const double sc = kssd_fold_scale(drfold);
	const double x = (double)xny * sc, s = (double)dsec * sc;
	const double m2 = (double)mut2 * sc, d = (double)dobj * sc;
	if (x <= 0.0)
		return 1;
	if (d <= 0.0)
		return 0;
	double dist = kssd_lm_core(x, s, m2, d, 0);
	if (dist < 0.05)
		dist = kssd_lm_core(x, s, m2, d, 1);
	if (dist < 0)
		dist = 0;
	return dist;
}

KSSD_HD double kssd_af_anib(uint32_t xny, uint32_t dobj, uint32_t dsec, uint32_t mut2,
			    uint32_t xctx, int drfold)
{
	const double sc = kssd_fold_scale(drfold);
	const double x = (double)xny * sc, s = (double)dsec * sc;
	const double m2 = (double)mut2 * sc, d = (double)dobj * sc, X = (double)xctx * sc;
	const double *p = kssd_p_af, *c = kssd_c_af;
	const double d1 = 1 / (p[0] * x + p[1] * s + p[2] * m2 + p[3] * d + p[4] * X + KSSD_EPS);
	const double d2 = 1 / (p[5] * x + p[6] * s + p[7] * m2 + p[8] * d + p[9] * X + KSSD_EPS);
	const double d3 = 1 / (p[10] * x + p[11] * s + p[12] * m2 + p[13] * d + p[14] * X + KSSD_EPS);
	return c[0] + c[1] * x + c[2] * s + c[3] * m2 + c[4] * d + c[5] * X + c[6] * d1 +
	       c[7] * d2 + c[8] * d3 + c[9] * x * d1 + c[10] * s * d1 + c[11] * m2 * d1 +
	       c[12] * d * d1 + c[13] * x * d2 + c[14] * s * d2 + c[15] * m2 * d2 +
	       c[16] * d * d2 + c[17] * x * d3 + c[18] * s * d3 + c[19] * m2 * d3 + c[20] * d * d3;
}

/* X = reference contexts, Y = query contexts, as the reference passes them
 * (command_ani.c:1339-1340); unsigned arithmetic as in the macros. */
KSSD_HD double kssd_mashD(uint32_t k, uint32_t X, uint32_t Y, uint32_t XnY)
{
	const double j = (double)XnY / (double)(uint32_t)(X + Y - XnY);
	return -log(2 * j / (1 + j)) / (double)k;
}

KSSD_HD double kssd_aafD(uint32_t k, uint32_t X, uint32_t Y, uint32_t XnY)
{
	const uint32_t mn = X < Y ? X : Y;
	return -log((double)XnY / (double)mn) / (double)k;
}

#endif
