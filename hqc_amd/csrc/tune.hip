// Launch parameters that the autotuner (hqc_amd/tune.py) may set per host thread.
// Every entry point reads them through hqc_tuned(key, default); 0 keeps the
// built-in default.  The table is thread-local so that host threads driving
// different streams (concurrent SCF buckets) keep their own choices.
#include "common.h"

static thread_local int g_tune[HQC_TUNE_COUNT];

int hqc_tuned(int key, int dflt) {
  const int v = (key >= 0 && key < HQC_TUNE_COUNT) ? g_tune[key] : 0;
  return v > 0 ? v : dflt;
}

extern "C" int hqc_set_tuning(int key, int value) {
  if (key < 0 || key >= HQC_TUNE_COUNT || value < 0) return HQC_E_ARG;
  if (key == HQC_TUNE_VMAT_TPW && value > 8) return HQC_E_ARG;
  if (key == HQC_TUNE_FFT_PLANE_THREADS && value != 0 && value != 256 && value != 512 &&
      value != 1024)
    return HQC_E_ARG;
  if ((key == HQC_TUNE_FFT_TWO_BUFFER && value > 1) || (key == HQC_TUNE_FFT_IP_ITEMS && value > 2) ||
      (key == HQC_TUNE_EIGH_RECOMPUTE && value > 1) ||
      (key == HQC_TUNE_FORCE_POINTWISE && value > 1) ||
      (key == HQC_TUNE_AO_POINTWISE && value > 1))
    return HQC_E_ARG;
  g_tune[key] = value;
  return HQC_OK;
}

extern "C" int hqc_get_tuning(int key) {
  if (key < 0 || key >= HQC_TUNE_COUNT) return -1;
  return g_tune[key];
}
