"""Oracle of the keyframe / source selection driver -- TEST INFRASTRUCTURE ONLY.

A plain-Python restatement of the reference node's message buffers and processImg
(evaluate.cc), used to check the library's ``pm::FrameDriver`` plan:

- ``GrabImage`` / ``GrabPose`` push the newest message to the FRONT of a list (:75-110);
- ``lastSize`` starts at aggNum * interval (:351);
- processImg (:137-276):
  - returns unless the image list is longer than lastSize (:157-160) and the newest pose is
    not older than the newest image (:170-172);
  - takes the newest image as the reference and looks for the pose with exactly its
    timestamp, walking the pose list from the front (:180-201); without one it returns
    (:197-201);
  - for i < aggNum it steps both iterators `interval` places further (:209-213): the
    sources are images and poses by list position;
  - then lastSize++ and both lists drop their oldest element (:234-237).

The reference dereferences past the end of the pose list when the source poses are not
buffered yet; the restatement (like the library) does not run such a step and counts it as
skipped.  Messages carry their grab-order index so the plan can be compared index by index.
"""
from __future__ import annotations


class DriverOracle:
    def __init__(self, agg_num: int = 3, interval: int = 6):
        self.agg_num = agg_num
        self.interval = interval
        self.images = []        # newest first: (stamp, index)
        self.poses = []         # newest first: (stamp, index)
        self.last_size = agg_num * interval
        self.n_images = 0
        self.n_poses = 0
        self.skipped = 0

    def grab_image(self, stamp: int):
        self.images.insert(0, (stamp, self.n_images))
        self.n_images += 1

    def grab_pose(self, stamp: int):
        self.poses.insert(0, (stamp, self.n_poses))
        self.n_poses += 1

    def process(self):
        """None, or (stamp, ref image, ref pose, [source images], [source poses])."""
        if len(self.images) <= self.last_size or not self.poses:
            return None
        t_img, ref_img = self.images[0]
        if self.poses[0][0] < t_img:
            return None
        k = next((k for k, (t, _) in enumerate(self.poses) if t == t_img), None)
        if k is None:
            return None
        reach = self.agg_num * self.interval
        if k + reach >= len(self.poses):
            self.skipped += 1
            return None
        src_img = [self.images[(i + 1) * self.interval][1] for i in range(self.agg_num)]
        src_pose = [self.poses[k + (i + 1) * self.interval][1] for i in range(self.agg_num)]
        ref_pose = self.poses[k][1]
        self.last_size += 1
        self.images.pop()
        self.poses.pop()
        return t_img, ref_img, ref_pose, src_img, src_pose
