// kNN kernels of joint dimension class 5 (see enm_knn.h).
#define ENM_KNN_KERNELS
#include "enm_knn.h"

ENM_KNN_DEFINE(5)
