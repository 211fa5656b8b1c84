// Package prtgpu binds the MI355X path-tracing core (libprt.so, C ABI in include/prt.h) to the
// renderer's Go interfaces: a batched geometry.Shape over the device BVH (shape.go), scene upload
// (scene.go) and a ProgressiveRaytracer with RenderPass / RenderProgressive (render.go).
//
// The package is meant to live inside the renderer's module as pkg/prtgpu, next to a copy of (or a
// symbolic link to) the library's include/ directory and libprt.so:
//
//	pkg/prtgpu/include/prt.h ...   pkg/prtgpu/libprt.so
//
// No Go pointer is retained by the library: every array passed in is copied during the call.
package prtgpu

/*
#cgo CFLAGS: -I${SRCDIR}/include
#cgo LDFLAGS: -L${SRCDIR} -lprt -Wl,-rpath,${SRCDIR}
#include <stdlib.h>
#include "prt.h"
*/
import "C"

import (
	"errors"
	"fmt"
	"unsafe"
)

// Error is a failed library call: its PRT_ERR_* code and the thread's last error message.
type Error struct {
	Code    int
	Message string
}

func (e *Error) Error() string { return fmt.Sprintf("prt error %d: %s", e.Code, e.Message) }

// ErrNoDevice is returned when no gfx950 device is visible.
var ErrNoDevice = errors.New("prtgpu: no usable GPU")

func check(rc C.int) error {
	if rc == C.PRT_OK {
		return nil
	}
	if rc == C.PRT_ERR_NO_DEVICE {
		return ErrNoDevice
	}
	return &Error{Code: int(rc), Message: C.GoString(C.prt_last_error())}
}

// DeviceCount returns the number of visible GPUs.
func DeviceCount() (int, error) {
	var n C.int32_t
	if err := check(C.prt_device_count(&n)); err != nil {
		return 0, err
	}
	return int(n), nil
}

// Ray and Hit are the 32-byte and 16-byte records of the batched intersection calls.
type Ray = C.prt_ray
type Hit = C.prt_hit

// deviceScene is a scene resident on one GPU (BVH, primitives, materials, lights).
type deviceScene struct {
	h *C.prt_scene
}

func (s *deviceScene) close() error {
	if s.h == nil {
		return nil
	}
	err := check(C.prt_scene_destroy(s.h)) // fails while a film still uses the scene
	if err == nil {
		s.h = nil
	}
	return err
}

// closestBatch fills hits[i] with the closest hit of rays[i] in [tmin, tmax] (prim_id -1 on a miss).
func (s *deviceScene) closestBatch(rays []Ray, hits []Hit) error {
	if len(rays) == 0 {
		return nil
	}
	if len(hits) < len(rays) {
		return errors.New("prtgpu: hit buffer shorter than the ray batch")
	}
	return check(C.prt_intersect_closest(s.h, &rays[0], C.int64_t(len(rays)), &hits[0], C.PRT_BUFFERS_HOST, nil))
}

// occludedBatch sets occluded[i] to 1 when some primitive lies on rays[i] within [tmin, tmax].
func (s *deviceScene) occludedBatch(rays []Ray, occluded []uint8) error {
	if len(rays) == 0 {
		return nil
	}
	if len(occluded) < len(rays) {
		return errors.New("prtgpu: occlusion buffer shorter than the ray batch")
	}
	return check(C.prt_intersect_any(s.h, &rays[0], C.int64_t(len(rays)),
		(*C.uint8_t)(unsafe.Pointer(&occluded[0])), C.PRT_BUFFERS_HOST, nil))
}
