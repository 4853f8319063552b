/*
 * CIE 1931 2-degree colour-matching functions, 380-700 nm in 5 nm steps
 * (65 samples) plus their 705-780 nm rows, and Johnson & Christy (1972) complex IOR of gold, copper and
 * silver, 380-680 nm in 25 nm steps.  These are published data tables; the
 * values are the ones rustracer uses (src/spectrum.rs:13-62 and 218-225) so
 * that spectral weights agree with the reference.  Data only, no code.
 */
#ifndef RT_SPECTRAL_DATA_H
#define RT_SPECTRAL_DATA_H
#define RT_CIE_N 65
#define RT_CIE_START 380.0f
#define RT_CIE_STEP 5.0f
#define RT_CIE_X_DATA \
    0.001368f, 0.002236f, 0.004243f, 0.007650f, 0.014310f, \
    0.023190f, 0.043510f, 0.077630f, 0.134380f, 0.214770f, \
    0.283900f, 0.328500f, 0.348280f, 0.348060f, 0.336200f, \
    0.318700f, 0.290800f, 0.251100f, 0.195360f, 0.142100f, \
    0.095640f, 0.057950f, 0.032010f, 0.014700f, 0.004900f, \
    0.002400f, 0.009300f, 0.029100f, 0.063270f, 0.109600f, \
    0.165500f, 0.225750f, 0.290400f, 0.359700f, 0.433450f, \
    0.512050f, 0.594500f, 0.678400f, 0.762100f, 0.842500f, \
    0.916300f, 0.978600f, 1.026300f, 1.056700f, 1.062200f, \
    1.045600f, 1.002600f, 0.938400f, 0.854450f, 0.751400f, \
    0.642400f, 0.541900f, 0.447900f, 0.360800f, 0.283500f, \
    0.218700f, 0.164900f, 0.121200f, 0.087400f, 0.063600f, \
    0.046770f, 0.032900f, 0.022700f, 0.015840f, 0.011359f
static const float RT_CIE_X[65] = {RT_CIE_X_DATA};

#define RT_CIE_Y_DATA \
    0.000039f, 0.000064f, 0.000120f, 0.000217f, 0.000396f, \
    0.000640f, 0.001210f, 0.002180f, 0.004000f, 0.007300f, \
    0.011600f, 0.016840f, 0.023000f, 0.029800f, 0.038000f, \
    0.048000f, 0.060000f, 0.073900f, 0.090980f, 0.112600f, \
    0.139020f, 0.169300f, 0.208020f, 0.258600f, 0.323000f, \
    0.407300f, 0.503000f, 0.608200f, 0.710000f, 0.793200f, \
    0.862000f, 0.914850f, 0.954000f, 0.980300f, 0.994950f, \
    1.000000f, 0.995000f, 0.978600f, 0.952000f, 0.915400f, \
    0.870000f, 0.816300f, 0.757000f, 0.694900f, 0.631000f, \
    0.566800f, 0.503000f, 0.441200f, 0.381000f, 0.321000f, \
    0.265000f, 0.217000f, 0.175000f, 0.138200f, 0.107000f, \
    0.081600f, 0.061000f, 0.044580f, 0.032000f, 0.023200f, \
    0.017000f, 0.011920f, 0.008210f, 0.005723f, 0.004102f
static const float RT_CIE_Y[65] = {RT_CIE_Y_DATA};

#define RT_CIE_Z_DATA \
    0.006450f, 0.010550f, 0.020050f, 0.036210f, 0.067850f, \
    0.110200f, 0.207400f, 0.371300f, 0.645600f, 1.039050f, \
    1.385600f, 1.622960f, 1.747060f, 1.782600f, 1.772110f, \
    1.744100f, 1.669200f, 1.528100f, 1.287640f, 1.041900f, \
    0.812950f, 0.616200f, 0.465180f, 0.353300f, 0.272000f, \
    0.212300f, 0.158200f, 0.111700f, 0.078250f, 0.057250f, \
    0.042160f, 0.029840f, 0.020300f, 0.013400f, 0.008750f, \
    0.005750f, 0.003900f, 0.002750f, 0.002100f, 0.001800f, \
    0.001650f, 0.001400f, 0.001100f, 0.001000f, 0.000800f, \
    0.000600f, 0.000340f, 0.000240f, 0.000190f, 0.000100f, \
    0.000050f, 0.000030f, 0.000020f, 0.000010f, 0.000000f, \
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, \
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f
static const float RT_CIE_Z[65] = {RT_CIE_Z_DATA};

/* The same observer's 705-780 nm rows (16 samples; z-bar is 0 there), for the
 * wide wavelength range [380, 780) of the four-wavelength mode (extension:
 * the reference's table stops at 700 nm).  CIE 1931 standard tabulation; its
 * chromaticity over these rows is the constant (0.734690, 0.265310). */
#define RT_CIE_WIDE_N 81
#define RT_CIE_X_TAIL_DATA \
    0.008110916f, 0.005790346f, 0.004109457f, 0.002899327f, 0.002049190f, \
    0.001439971f, 0.000999949f, 0.000690079f, 0.000476021f, 0.000332301f, \
    0.000234826f, 0.000166151f, 0.000117413f, 0.0000830753f, 0.0000587065f, \
    0.0000415099f
#define RT_CIE_Y_TAIL_DATA \
    0.002929000f, 0.002091000f, 0.001484000f, 0.001047000f, 0.000740000f, \
    0.000520000f, 0.000361100f, 0.000249200f, 0.000171900f, 0.000120000f, \
    0.000084800f, 0.000060000f, 0.000042400f, 0.0000300000f, 0.0000212000f, \
    0.0000149900f
#define RT_CIE_Z_TAIL_DATA \
    0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f
static const float RT_CIE_WX[81] = {RT_CIE_X_DATA, RT_CIE_X_TAIL_DATA};
static const float RT_CIE_WY[81] = {RT_CIE_Y_DATA, RT_CIE_Y_TAIL_DATA};
static const float RT_CIE_WZ[81] = {RT_CIE_Z_DATA, RT_CIE_Z_TAIL_DATA};
#define RT_GOLD_N_DATA 1.69f, 1.65f, 1.54f, 1.33f, 0.97f, 0.54f, 0.30f, 0.23f, 0.20f, 0.19f, 0.18f, 0.16f, 0.16f
static const float RT_GOLD_N[13] = {RT_GOLD_N_DATA};
#define RT_GOLD_K_DATA 1.91f, 1.94f, 1.95f, 1.92f, 1.80f, 1.93f, 2.41f, 3.00f, 3.56f, 4.07f, 4.56f, 4.99f, 5.41f
static const float RT_GOLD_K[13] = {RT_GOLD_K_DATA};
#define RT_COPPER_N_DATA 1.23f, 1.21f, 1.16f, 1.10f, 1.04f, 0.88f, 0.60f, 0.38f, 0.26f, 0.22f, 0.21f, 0.20f, 0.21f
static const float RT_COPPER_N[13] = {RT_COPPER_N_DATA};
#define RT_COPPER_K_DATA 2.00f, 2.26f, 2.49f, 2.59f, 2.64f, 2.69f, 2.84f, 3.21f, 3.70f, 4.17f, 4.64f, 5.07f, 5.48f
static const float RT_COPPER_K[13] = {RT_COPPER_K_DATA};
#define RT_SILVER_N_DATA 0.18f, 0.14f, 0.12f, 0.13f, 0.13f, 0.13f, 0.14f, 0.14f, 0.15f, 0.16f, 0.16f, 0.17f, 0.17f
static const float RT_SILVER_N[13] = {RT_SILVER_N_DATA};
#define RT_SILVER_K_DATA 1.57f, 2.01f, 2.50f, 3.02f, 3.55f, 4.05f, 4.54f, 5.02f, 5.50f, 5.98f, 6.42f, 6.84f, 7.22f
static const float RT_SILVER_K[13] = {RT_SILVER_K_DATA};
#define RT_METAL_N 13
#define RT_METAL_START 380.0f
#define RT_METAL_STEP 25.0f
#endif
