// Upload-once frame ring: device-resident ingested frames for StereoMapper.
//
// The reference uploads and converts a frame every time it is used (StereoMapper::update,
// stereo_mapper.cpp:73-77: host u8 -> float, then a host-to-device copy per call), and its
// ROS node uses each frame up to aggNum + 1 times (once as the reference, then as a source
// of aggNum later keyframes, evaluate.cc:209-233).  A FrameRing holds the newest `capacity`
// frames on the device, each uploaded once (through pinned staging) and converted once into
// the cost kernel's quad texture; StereoMapper's frame-id overloads of initReference and
// update read them from there.
//
// Frame f lives in slot f % capacity until another frame is put into that slot.  put()
// enqueues the upload and conversion on the given stream and records an event; the mapper
// waits for that event on its own stream.  A frame staged by StereoMapper::update must stay
// in the ring until the mapper's next output call returns (the mapper checks this and raises
// PM_ERR_INVALID_STATE otherwise).
#ifndef PM_FRAME_RING_H
#define PM_FRAME_RING_H

#include <cstdint>
#include <vector>

#include "pm/types.h"
#include "pm_capi.h"

namespace pm {

class FrameRing {
 public:
  FrameRing(int width, int height, int capacity);
  ~FrameRing();
  FrameRing(const FrameRing&) = delete;
  FrameRing& operator=(const FrameRing&) = delete;

  // Uploads an 8-bit grayscale width x height image as frame `frame` (>= 0) and builds its
  // quad texture, enqueued on `stream` (a hipStream_t; null = the default stream).  The
  // host image may be reused as soon as put returns.
  void put(int64_t frame, const Mat& image, void* stream);
  bool holds(int64_t frame) const;
  // device pointers of a held frame (PM_ERR_INVALID_STATE if it is not held)
  const uint8_t* image(int64_t frame) const;    // u8 [H][W]
  const uint64_t* quad(int64_t frame) const;    // 8-byte texels [H][W]
  void* ready_event(int64_t frame) const;       // hipEvent_t recorded after the conversion

  int width() const { return W_; }
  int height() const { return H_; }
  int capacity() const { return (int)slots_.size(); }
  int64_t puts() const { return puts_; }         // frames uploaded so far

 private:
  struct Slot {
    int64_t frame = -1;
    DeviceMat img, quad;
    uint8_t* pinned = nullptr;
    void* event = nullptr;
    bool pending = false;
  };
  const Slot& held(int64_t frame) const;
  void release();

  int W_, H_;
  std::vector<Slot> slots_;
  int64_t puts_ = 0;
};

// the FrameRing behind a C handle (null for a null handle)
FrameRing* ring_of(const pm_frame_ring* r);

}  // namespace pm

#endif  // PM_FRAME_RING_H
