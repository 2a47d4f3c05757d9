// The joint advection kernel (k_adv_t: nut, the wind advection and the moisture advection in one launch) in a
// compilation of its own: k_step.hip with OBL_ADV_ONLY keeps that kernel and its launcher and leaves the rest out,
// so that the build can give it its own instruction scheduling (the maximum-ILP strategy, which shortens it by ~3 %
// and lengthens the physics kernel, so it is not used for the other kernels).
#define OBL_ADV_ONLY
#include "k_step.hip"
