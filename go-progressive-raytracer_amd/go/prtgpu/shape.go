package prtgpu

/*
#include "prt.h"
*/
import "C"

import (
	"math"

	"github.com/df07/go-progressive-raytracer/pkg/core"
	"github.com/df07/go-progressive-raytracer/pkg/geometry"
	"github.com/df07/go-progressive-raytracer/pkg/material"
)

// BVHShape is the device BVH as a geometry.Shape (pkg/geometry/interfaces.go:9-12), the seam at which
// *geometry.BVH (bvh.go:183-239) is replaced. The library answers whole batches; HitBatch is the call
// a batching caller uses (one batch per tile and bounce), and Hit serves the per-ray interface with a
// batch of one.
//
// A device hit names the closest primitive and its distance; the reference's SurfaceInteraction
// (point, face-forwarded normal, UV, material) comes from that primitive's own float64 Hit on the same
// ray, so the interaction is exactly the reference's. If the float32 traversal and the float64 test
// disagree at an edge (the primitive's own test misses), the ray is reported as a miss.
type BVHShape struct {
	s *Scene
}

// Shape returns the scene's device BVH as a geometry.Shape.
func (s *Scene) Shape() *BVHShape { return &BVHShape{s: s} }

func toRay(r core.Ray, tMin, tMax float64) Ray {
	var out Ray
	out.org[0], out.org[1], out.org[2] = C.float(r.Origin.X), C.float(r.Origin.Y), C.float(r.Origin.Z)
	out.dir[0], out.dir[1], out.dir[2] = C.float(r.Direction.X), C.float(r.Direction.Y), C.float(r.Direction.Z)
	out.tmin = C.float(tMin)
	if math.IsInf(tMax, 1) || tMax > math.MaxFloat32 {
		out.tmax = C.float(math.Inf(1))
	} else {
		out.tmax = C.float(tMax)
	}
	return out
}

// HitBatch intersects rays[i] on [tMin[i], tMax[i]] and returns, per ray, the surface interaction of the
// closest hit or nil.
func (b *BVHShape) HitBatch(rays []core.Ray, tMin, tMax []float64) ([]*material.SurfaceInteraction, error) {
	rs := make([]Ray, len(rays))
	for i, r := range rays {
		rs[i] = toRay(r, tMin[i], tMax[i])
	}
	hits := make([]Hit, len(rays))
	if err := b.s.dev.closestBatch(rs, hits); err != nil {
		return nil, err
	}
	out := make([]*material.SurfaceInteraction, len(rays))
	for i, h := range hits {
		if h.prim_id < 0 || int(h.prim_id) >= len(b.s.prims) {
			continue
		}
		if si, ok := b.s.prims[h.prim_id].Hit(rays[i], tMin[i], tMax[i]); ok {
			out[i] = si
		}
	}
	return out, nil
}

// OccludedBatch reports for every ray whether anything lies on [tMin[i], tMax[i]] (shadow rays; the
// reference traces them as closest hits, path_tracing.go:116-157).
func (b *BVHShape) OccludedBatch(rays []core.Ray, tMin, tMax []float64) ([]bool, error) {
	rs := make([]Ray, len(rays))
	for i, r := range rays {
		rs[i] = toRay(r, tMin[i], tMax[i])
	}
	occ := make([]uint8, len(rays))
	if err := b.s.dev.occludedBatch(rs, occ); err != nil {
		return nil, err
	}
	out := make([]bool, len(rays))
	for i, o := range occ {
		out[i] = o != 0
	}
	return out, nil
}

// Hit implements geometry.Shape for one ray (a batch of one: correct, but batching callers should use
// HitBatch; a GPU needs a million rays in flight).
func (b *BVHShape) Hit(ray core.Ray, tMin, tMax float64) (*material.SurfaceInteraction, bool) {
	out, err := b.HitBatch([]core.Ray{ray}, []float64{tMin}, []float64{tMax})
	if err != nil || out[0] == nil {
		return nil, false
	}
	return out[0], true
}

// BoundingBox implements geometry.Shape: the union of the scene's shape boxes (the reference BVH root).
func (b *BVHShape) BoundingBox() geometry.AABB { return b.s.bbox }

var _ geometry.Shape = (*BVHShape)(nil)
