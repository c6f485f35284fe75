// Upload-once frame ring (include/pm/frame_ring.h) and its C ABI.
// Replaces the per-use upload and conversion of StereoMapper::update (stereo_mapper.cpp:73-77).
#include "pm/frame_ring.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

#include "common.hpp"
#include "pm/stereo_mapper.h"

namespace pm {

namespace {

void ring_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(e == hipErrorOutOfMemory ? PM_ERR_OUT_OF_MEMORY : PM_ERR_DEVICE,
                std::string("FrameRing: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

FrameRing::FrameRing(int width, int height, int capacity) : W_(width), H_(height) {
  if (width <= 0 || height <= 0 || width > 65535 || height > 65535)
    throw Error(PM_ERR_INVALID_ARGUMENT, "FrameRing: bad image size");
  if (capacity < 1 || capacity > 4096)
    throw Error(PM_ERR_INVALID_ARGUMENT, "FrameRing: capacity must be in 1..4096");
  slots_.resize(capacity);
  const size_t N = (size_t)W_ * H_;
  try {
    for (Slot& s : slots_) {
      void* p = nullptr;
      ring_check(hipMalloc(&p, N), "hipMalloc");
      s.img.rows = H_; s.img.cols = W_; s.img.type = CV_8U; s.img.step = (size_t)W_;
      s.img.data = static_cast<uint8_t*>(p);
      p = nullptr;
      ring_check(hipMalloc(&p, N * 8), "hipMalloc");
      s.quad.rows = H_; s.quad.cols = 2 * W_; s.quad.type = CV_32S; s.quad.step = (size_t)W_ * 8;
      s.quad.data = static_cast<uint8_t*>(p);
      p = nullptr;
      ring_check(hipHostMalloc(&p, N, hipHostMallocDefault), "hipHostMalloc");
      s.pinned = static_cast<uint8_t*>(p);
      hipEvent_t ev = nullptr;
      ring_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
      s.event = ev;
    }
  } catch (...) {
    release();
    throw;
  }
}

FrameRing::~FrameRing() { release(); }

void FrameRing::release() {
  for (Slot& s : slots_) {
    if (s.event) {
      if (s.pending) (void)hipEventSynchronize((hipEvent_t)s.event);
      (void)hipEventDestroy((hipEvent_t)s.event);
    }
    if (s.pinned) (void)hipHostFree(s.pinned);
    if (s.img.data) (void)hipFree(s.img.data);
    if (s.quad.data) (void)hipFree(s.quad.data);
    s = Slot();
  }
}

void FrameRing::put(int64_t frame, const Mat& image, void* stream) {
  if (frame < 0) throw Error(PM_ERR_INVALID_ARGUMENT, "FrameRing::put: frame ids are >= 0");
  if (image.empty() || image.type != CV_8U || image.rows != H_ || image.cols != W_)
    throw Error(PM_ERR_INVALID_ARGUMENT, "FrameRing::put: image must be 8-bit grayscale " +
                                             std::to_string(W_) + "x" + std::to_string(H_));
  Slot& s = slots_[(size_t)(frame % (int64_t)slots_.size())];
  // the staging buffer may still feed the slot's previous upload
  if (s.pending) ring_check(hipEventSynchronize((hipEvent_t)s.event), "hipEventSynchronize");
  s.pending = false;
  s.frame = -1;
  for (int y = 0; y < H_; ++y)
    std::memcpy(s.pinned + (size_t)y * W_, image.data + (size_t)y * image.step, (size_t)W_);
  hipStream_t st = (hipStream_t)stream;
  ring_check(hipMemcpyAsync(s.img.data, s.pinned, (size_t)W_ * H_, hipMemcpyHostToDevice, st),
             "upload");
  if (pm_make_quad(s.img.data, reinterpret_cast<uint64_t*>(s.quad.data), W_, H_, stream) != PM_OK)
    throw Error(PM_ERR_LAUNCH, std::string("FrameRing::put: ") + pm_last_error());
  ring_check(hipEventRecord((hipEvent_t)s.event, st), "hipEventRecord");
  s.pending = true;
  s.frame = frame;
  ++puts_;
}

bool FrameRing::holds(int64_t frame) const {
  return frame >= 0 && slots_[(size_t)(frame % (int64_t)slots_.size())].frame == frame;
}

const FrameRing::Slot& FrameRing::held(int64_t frame) const {
  if (!holds(frame))
    throw Error(PM_ERR_INVALID_STATE, "FrameRing: frame " + std::to_string(frame) +
                                          " is not in the ring (never put, or evicted)");
  return slots_[(size_t)(frame % (int64_t)slots_.size())];
}

const uint8_t* FrameRing::image(int64_t frame) const { return held(frame).img.data; }

const uint64_t* FrameRing::quad(int64_t frame) const {
  return reinterpret_cast<const uint64_t*>(held(frame).quad.data);
}

void* FrameRing::ready_event(int64_t frame) const { return held(frame).event; }

}  // namespace pm

// ---- C wrappers ---------------------------------------------------------------------------

struct pm_frame_ring {
  pm::FrameRing* impl;
};

namespace {

template <typename F>
pm_status ring_guard(F&& f) {
  pm::clear_error();
  try {
    f();
    return PM_OK;
  } catch (const pm::Error& e) {
    return pm::fail(e.status, e.what());
  } catch (const std::bad_alloc&) {
    return pm::fail(PM_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return pm::fail(PM_ERR_DEVICE, e.what());
  }
}

}  // namespace

extern "C" pm_status pm_frame_ring_create(int32_t width, int32_t height, int32_t capacity,
                                          pm_frame_ring** out) {
  if (!out) return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_frame_ring_create: null argument");
  *out = nullptr;
  return ring_guard([&] {
    auto* r = new pm_frame_ring{nullptr};
    try {
      r->impl = new pm::FrameRing(width, height, capacity);
    } catch (...) {
      delete r;
      throw;
    }
    *out = r;
  });
}

extern "C" void pm_frame_ring_destroy(pm_frame_ring* r) {
  if (!r) return;
  delete r->impl;
  delete r;
}

extern "C" pm_status pm_frame_ring_put(pm_frame_ring* r, int64_t frame, const uint8_t* image,
                                       pm_stream stream) {
  if (!r || !r->impl || !image)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_frame_ring_put: null argument");
  return ring_guard([&] {
    r->impl->put(frame, pm::Mat(r->impl->height(), r->impl->width(), pm::CV_8U,
                                const_cast<uint8_t*>(image)),
                 stream);
  });
}

extern "C" pm_status pm_frame_ring_frame(const pm_frame_ring* r, int64_t frame,
                                         const uint8_t** image, const uint64_t** quad) {
  if (!r || !r->impl) return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_frame_ring_frame: null ring");
  return ring_guard([&] {
    if (image) *image = r->impl->image(frame);
    if (quad) *quad = r->impl->quad(frame);
  });
}

extern "C" pm_status pm_frame_ring_puts(const pm_frame_ring* r, int64_t* puts) {
  if (!r || !r->impl || !puts)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_frame_ring_puts: null argument");
  *puts = r->impl->puts();
  return PM_OK;
}

namespace pm {
FrameRing* ring_of(const pm_frame_ring* r) { return r ? r->impl : nullptr; }
}  // namespace pm
