// Frame clock of the streaming ring (ddm_clock_* in include/ddm.h): maps live
// capture times to history slots, and turns what the slots hold into the
// (order, mask) pair of a temporal pass.  Host bookkeeping only.
//
// The contract is the reference's FrameClock (motility-ddm
// src/camera.ts:306-425): frame n = round((t - t0) / dt) lives in slot n mod T;
// a time index of a snapshot is valid only while its slot still holds that
// very frame, so dropped, late, duplicated and overwritten frames drop out of
// the mask by construction.  This implementation keeps, per slot, the frame
// index it holds; the window's validity is then "slot value inside the
// window", one comparison per slot.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "../../include/ddm.h"

struct ddm_clock {
  int32_t T = 0;
  double period_ms = 0;
  // the current epoch: its first time stamp and timeline
  bool anchored = false;
  double anchor_ms = 0;
  int32_t source = -1;      // DDM_TIME_* of the last placed frame; -1 in a fresh epoch
  int64_t newest = -1;      // newest frame index placed in this epoch
  std::vector<int64_t> held;   // frame index held by each slot, -1 for none
  // statistics (kept across epochs)
  int64_t placed = 0, missed = 0, duplicates = 0, restarts = 0;
  bool have_prev = false;   // the previous placed frame's time stamp, within the epoch
  double prev_ms = 0;
  static constexpr int kGaps = 120;   // intervals behind the median
  std::array<double, kGaps> gaps{};
  int ngaps = 0, gap_next = 0;
};

namespace {

bool media(int32_t src) { return src == DDM_TIME_MEDIA; }

// Math.round: the nearest integer, halves toward +infinity (exact: the
// fractional part x - floor(x) is computed without rounding)
int64_t round_half_up(double x) {
  const double f = std::floor(x);
  return (int64_t)f + ((x - f) >= 0.5 ? 1 : 0);
}

void new_epoch(ddm_clock* c) {
  c->anchored = false;
  c->source = -1;
  c->newest = -1;
  std::fill(c->held.begin(), c->held.end(), (int64_t)-1);
  c->have_prev = false;
  c->restarts++;
}

int64_t mod_t(int64_t n, int32_t T) { return ((n % T) + T) % T; }

}  // namespace

extern "C" {

int ddm_clock_create(double fps, int32_t T, ddm_clock** out) {
  if (!out) return DDM_EINVAL;
  *out = nullptr;
  if (!(fps > 0) || !std::isfinite(fps) || T < 1) return DDM_EINVAL;
  ddm_clock* c = new (std::nothrow) ddm_clock();
  if (!c) return DDM_ENOMEM;
  c->T = T;
  c->period_ms = 1000.0 / fps;
  c->held.assign(T, -1);
  *out = c;
  return DDM_OK;
}

void ddm_clock_destroy(ddm_clock* c) { delete c; }

int ddm_clock_place(ddm_clock* c, double t_ms, int32_t source, int32_t* slot, int64_t* n_out) {
  if (!c || !std::isfinite(t_ms) || source < DDM_TIME_CAPTURE || source > DDM_TIME_MEDIA) return DDM_EINVAL;
  // the media timeline and the performance clock (capture, watchdog) do not mix
  if (c->source >= 0 && media(source) != media(c->source)) new_epoch(c);
  c->source = source;
  if (!c->anchored) {
    c->anchored = true;
    c->anchor_ms = t_ms;
  }
  int64_t n = round_half_up((t_ms - c->anchor_ms) / c->period_ms);
  if (n <= c->newest) {   // this frame, or a later one, is already placed
    c->duplicates++;
    return DDM_DROPPED;
  }
  if (c->newest >= 0 && n - c->newest > c->T) {
    // a gap longer than the ring leaves nothing usable: a new epoch at this frame
    new_epoch(c);
    c->source = source;
    c->anchored = true;
    c->anchor_ms = t_ms;
    n = 0;
  }
  if (c->newest >= 0) c->missed += n - c->newest - 1;
  if (c->have_prev) {
    c->gaps[c->gap_next] = t_ms - c->prev_ms;
    c->gap_next = (c->gap_next + 1) % ddm_clock::kGaps;
    c->ngaps = std::min(c->ngaps + 1, ddm_clock::kGaps);
  }
  c->have_prev = true;
  c->prev_ms = t_ms;
  const int64_t s = n % c->T;
  c->held[s] = n;
  c->newest = n;
  c->placed++;
  if (slot) *slot = (int32_t)s;
  if (n_out) *n_out = n;
  return DDM_OK;
}

int ddm_clock_drop(ddm_clock* c, int64_t n) {
  if (!c) return DDM_EINVAL;
  if (n >= 0 && c->held[n % c->T] == n) c->held[n % c->T] = -1;
  return DDM_OK;
}

int ddm_clock_reset(ddm_clock* c) {
  if (!c) return DDM_EINVAL;
  new_epoch(c);
  return DDM_OK;
}

int ddm_clock_state_get(const ddm_clock* c, ddm_clock_state* s) {
  if (!c || !s) return DDM_EINVAL;
  s->newest_n = c->newest;
  s->frames = c->placed;
  s->missed = c->missed;
  s->duplicates = c->duplicates;
  s->resets = c->restarts;
  s->time_source = c->source;
  // a slot's frame is in the window iff it is newer than newest - T
  const int64_t oldest = c->newest - c->T + 1;
  int32_t valid = 0;
  for (int64_t v : c->held) valid += (v >= 0 && v >= oldest) ? 1 : 0;
  s->valid = valid;
  if (c->ngaps == 0) {
    s->median_dt_ms = NAN;
  } else {
    std::array<double, ddm_clock::kGaps> tmp;
    std::copy(c->gaps.begin(), c->gaps.begin() + c->ngaps, tmp.begin());
    const int k = c->ngaps / 2;   // the upper median for an even count
    std::nth_element(tmp.begin(), tmp.begin() + k, tmp.begin() + c->ngaps);
    s->median_dt_ms = tmp[k];
  }
  return DDM_OK;
}

int ddm_clock_snapshot(const ddm_clock* c, int32_t guard, uint32_t* order, float* mask, int32_t* used) {
  if (!c || !order || !mask || guard < 0) return DDM_EINVAL;
  const int64_t oldest = c->newest - c->T + 1;   // frame of time index 0
  int32_t u = 0;
  for (int32_t j = 0; j < c->T; j++) {
    const int64_t n = oldest + j;
    const int64_t s = mod_t(n, c->T);
    const bool ok = n >= 0 && j >= guard && c->held[s] == n;
    order[j] = (uint32_t)s;
    mask[j] = ok ? 1.f : 0.f;
    u += ok ? 1 : 0;
  }
  if (used) *used = u;
  return DDM_OK;
}

int ddm_shuffle_order(uint32_t* order, float* mask, int32_t n, const double* u) {
  if (n < 0 || (n > 0 && (!order || !mask)) || (n > 1 && !u)) return DDM_EINVAL;
  for (int32_t k = 0; k + 1 < n; k++)
    if (!(u[k] >= 0 && u[k] < 1)) return DDM_EINVAL;
  // Durstenfeld's form of Fisher-Yates: position i = n-1 .. 1 trades places
  // with floor(u (i + 1)); draw k belongs to position n-1-k
  for (int32_t i = n - 1; i > 0; i--) {
    const int32_t j = (int32_t)std::floor(u[n - 1 - i] * (i + 1));
    std::swap(order[i], order[j]);
    std::swap(mask[i], mask[j]);
  }
  return DDM_OK;
}

}  // extern "C"
