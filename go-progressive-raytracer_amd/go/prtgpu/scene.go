package prtgpu

/*
#include <stdlib.h>
#include <string.h>
#include "prt.h"
*/
import "C"

import (
	"fmt"
	"unsafe"

	"github.com/df07/go-progressive-raytracer/pkg/core"
	"github.com/df07/go-progressive-raytracer/pkg/geometry"
	"github.com/df07/go-progressive-raytracer/pkg/lights"
	"github.com/df07/go-progressive-raytracer/pkg/material"
	"github.com/df07/go-progressive-raytracer/pkg/scene"
)

// Scene is a scene.Scene uploaded to one GPU (Scene.Preprocess, pkg/scene/scene.go:45-78, on the
// device: the library builds its own BVH). prims maps the library's primitive ids (prt.h: mesh
// triangles mesh by mesh, then spheres, quads, 6 quads per box) back to the Go shapes, so that a
// device hit becomes the reference's *material.SurfaceInteraction.
type Scene struct {
	dev   deviceScene
	src   *scene.Scene
	prims []geometry.Shape
	bbox  geometry.AABB
}

// cbuf is C memory for the descriptor's arrays (cgo forbids Go memory that holds Go pointers).
type cbuf struct{ ptrs []unsafe.Pointer }

func (b *cbuf) alloc(n, size int) unsafe.Pointer {
	if n == 0 {
		return nil
	}
	p := C.calloc(C.size_t(n), C.size_t(size))
	b.ptrs = append(b.ptrs, p)
	return p
}

func (b *cbuf) free() {
	for _, p := range b.ptrs {
		C.free(p)
	}
}

func v3(dst *[3]C.double, v core.Vec3) {
	dst[0], dst[1], dst[2] = C.double(v.X), C.double(v.Y), C.double(v.Z)
}

type flattener struct {
	mats     []C.prt_material
	matIndex map[material.Material]int
	textures []*material.ImageTexture
	spheres  []C.prt_sphere
	quads    []C.prt_quad
	boxes    []C.prt_box
	lights   []C.prt_light
	// meshes: triangles of each mesh as unshared vertices, with per-face materials and normals
	meshVerts [][]float64
	meshMats  [][]int32
	meshNorms [][]float64
	meshTris  [][]geometry.Shape
	sphereOf  map[*geometry.Sphere]int
	quadOf    map[*geometry.Quad]int
	sphereSh  []geometry.Shape
	quadSh    []geometry.Shape
	boxSh     []geometry.Shape
}

func (f *flattener) material(m material.Material) (int, error) {
	if i, ok := f.matIndex[m]; ok {
		return i, nil
	}
	var pm C.prt_material
	switch t := m.(type) {
	case *material.Lambertian:
		pm._type = C.PRT_MAT_LAMBERTIAN
		if err := f.albedo(&pm, t.Albedo); err != nil {
			return 0, err
		}
	case *material.Metal:
		pm._type = C.PRT_MAT_METAL
		pm.fuzz = C.double(t.Fuzzness)
		if err := f.albedo(&pm, t.Albedo); err != nil {
			return 0, err
		}
	case *material.Dielectric:
		pm._type = C.PRT_MAT_DIELECTRIC
		pm.ior = C.double(t.RefractiveIndex)
	case *material.Emissive:
		pm._type = C.PRT_MAT_EMISSIVE
		v3(&pm.emission, t.Emission)
	case *material.Layered:
		o, err := f.material(t.Outer)
		if err != nil {
			return 0, err
		}
		in, err := f.material(t.Inner)
		if err != nil {
			return 0, err
		}
		pm._type, pm.outer, pm.inner = C.PRT_MAT_LAYERED, C.int32_t(o), C.int32_t(in)
	case *material.Mix:
		a, err := f.material(t.Material1)
		if err != nil {
			return 0, err
		}
		b, err := f.material(t.Material2)
		if err != nil {
			return 0, err
		}
		pm._type, pm.outer, pm.inner, pm.ratio = C.PRT_MAT_MIX, C.int32_t(a), C.int32_t(b), C.double(t.Ratio)
	default:
		return 0, fmt.Errorf("prtgpu: material %T is not supported by the device path", m)
	}
	f.mats = append(f.mats, pm)
	f.matIndex[m] = len(f.mats) - 1
	return len(f.mats) - 1, nil
}

func (f *flattener) albedo(pm *C.prt_material, src material.ColorSource) error {
	switch c := src.(type) {
	case *material.SolidColor:
		v3(&pm.albedo, c.Color)
	case *material.ImageTexture:
		f.textures = append(f.textures, c)
		pm.texture = C.int32_t(len(f.textures))
	default:
		return fmt.Errorf("prtgpu: colour source %T is not supported by the device path", src)
	}
	return nil
}

func (f *flattener) shape(sh geometry.Shape) error {
	switch t := sh.(type) {
	case *geometry.Sphere:
		m, err := f.material(t.Material)
		if err != nil {
			return err
		}
		var ps C.prt_sphere
		v3(&ps.center, t.Center)
		ps.radius, ps.material = C.double(t.Radius), C.int32_t(m)
		f.sphereOf[t] = len(f.spheres)
		f.spheres = append(f.spheres, ps)
		f.sphereSh = append(f.sphereSh, t)
	case *geometry.Quad:
		m, err := f.material(t.Material)
		if err != nil {
			return err
		}
		var pq C.prt_quad
		v3(&pq.corner, t.Corner)
		v3(&pq.u, t.U)
		v3(&pq.v, t.V)
		pq.material = C.int32_t(m)
		f.quadOf[t] = len(f.quads)
		f.quads = append(f.quads, pq)
		f.quadSh = append(f.quadSh, t)
	case *geometry.Box:
		m, err := f.material(t.Material)
		if err != nil {
			return err
		}
		var pb C.prt_box
		v3(&pb.center, t.Center)
		v3(&pb.size, t.Size)
		v3(&pb.rotation, t.Rotation)
		pb.material = C.int32_t(m)
		f.boxes = append(f.boxes, pb)
		f.boxSh = append(f.boxSh, t)
	case *geometry.TriangleMesh:
		tris := t.GetTriangles()
		verts := make([]float64, 0, 9*len(tris))
		mats := make([]int32, 0, len(tris))
		norms := make([]float64, 0, 3*len(tris))
		for _, s := range tris {
			tr, ok := s.(*geometry.Triangle)
			if !ok {
				return fmt.Errorf("prtgpu: mesh element %T is not a triangle", s)
			}
			m, err := f.material(tr.Material)
			if err != nil {
				return err
			}
			n := tr.GetNormal()
			verts = append(verts, tr.V0.X, tr.V0.Y, tr.V0.Z, tr.V1.X, tr.V1.Y, tr.V1.Z, tr.V2.X, tr.V2.Y, tr.V2.Z)
			mats = append(mats, int32(m))
			norms = append(norms, n.X, n.Y, n.Z)
		}
		f.meshVerts = append(f.meshVerts, verts)
		f.meshMats = append(f.meshMats, mats)
		f.meshNorms = append(f.meshNorms, norms)
		f.meshTris = append(f.meshTris, tris)
	default:
		return fmt.Errorf("prtgpu: shape %T is not supported by this binding", sh)
	}
	return nil
}

func (f *flattener) light(l lights.Light) error {
	var pl C.prt_light
	pl.shape = -1
	switch t := l.(type) {
	case *lights.QuadLight:
		i, ok := f.quadOf[t.Quad]
		if !ok { // a light whose quad is not among the shapes still needs its shape on the device
			if err := f.shape(t.Quad); err != nil {
				return err
			}
			i = f.quadOf[t.Quad]
		}
		pl._type, pl.shape = C.PRT_LIGHT_QUAD, C.int32_t(i)
	case *lights.SphereLight:
		i, ok := f.sphereOf[t.Sphere]
		if !ok {
			if err := f.shape(t.Sphere); err != nil {
				return err
			}
			i = f.sphereOf[t.Sphere]
		}
		pl._type, pl.shape = C.PRT_LIGHT_SPHERE, C.int32_t(i)
	case *lights.UniformInfiniteLight:
		pl._type = C.PRT_LIGHT_UNIFORM_INFINITE
		v3(&pl.emission, t.Emit(core.NewRay(core.Vec3{}, core.NewVec3(0, 1, 0)), nil))
	case *lights.GradientInfiniteLight:
		// Emit lerps bottom -> top with 0.5 (dir.y + 1): the straight-up and straight-down rays give the ends
		pl._type = C.PRT_LIGHT_GRADIENT_INFINITE
		v3(&pl.top, t.Emit(core.NewRay(core.Vec3{}, core.NewVec3(0, 1, 0)), nil))
		v3(&pl.bottom, t.Emit(core.NewRay(core.Vec3{}, core.NewVec3(0, -1, 0)), nil))
	default:
		return fmt.Errorf("prtgpu: light %T is not supported by this binding", l)
	}
	f.lights = append(f.lights, pl)
	return nil
}

// NewScene flattens s (shapes, lights, materials, camera) and uploads it to GPU `device`.
func NewScene(device int, s *scene.Scene) (*Scene, error) {
	f := &flattener{matIndex: map[material.Material]int{}, sphereOf: map[*geometry.Sphere]int{},
		quadOf: map[*geometry.Quad]int{}}
	for _, sh := range s.Shapes {
		if err := f.shape(sh); err != nil {
			return nil, err
		}
	}
	for _, l := range s.Lights {
		if err := f.light(l); err != nil {
			return nil, err
		}
	}
	var b cbuf
	defer b.free()
	var d C.prt_scene_desc
	copyTo := func(n, size int, src unsafe.Pointer) unsafe.Pointer {
		p := b.alloc(n, size)
		if n > 0 {
			C.memcpy(p, src, C.size_t(n*size))
		}
		return p
	}
	if len(f.mats) > 0 {
		d.materials = (*C.prt_material)(copyTo(len(f.mats), int(unsafe.Sizeof(f.mats[0])), unsafe.Pointer(&f.mats[0])))
	}
	d.num_materials = C.int32_t(len(f.mats))
	if len(f.spheres) > 0 {
		d.spheres = (*C.prt_sphere)(copyTo(len(f.spheres), int(unsafe.Sizeof(f.spheres[0])), unsafe.Pointer(&f.spheres[0])))
	}
	d.num_spheres = C.int32_t(len(f.spheres))
	if len(f.quads) > 0 {
		d.quads = (*C.prt_quad)(copyTo(len(f.quads), int(unsafe.Sizeof(f.quads[0])), unsafe.Pointer(&f.quads[0])))
	}
	d.num_quads = C.int32_t(len(f.quads))
	if len(f.boxes) > 0 {
		d.boxes = (*C.prt_box)(copyTo(len(f.boxes), int(unsafe.Sizeof(f.boxes[0])), unsafe.Pointer(&f.boxes[0])))
	}
	d.num_boxes = C.int32_t(len(f.boxes))
	if len(f.lights) > 0 {
		d.lights = (*C.prt_light)(copyTo(len(f.lights), int(unsafe.Sizeof(f.lights[0])), unsafe.Pointer(&f.lights[0])))
	}
	d.num_lights = C.int32_t(len(f.lights))
	if n := len(f.meshVerts); n > 0 {
		meshes := unsafe.Slice((*C.prt_mesh)(b.alloc(n, int(unsafe.Sizeof(C.prt_mesh{})))), n)
		for k := range f.meshVerts {
			nt := len(f.meshMats[k])
			faces := unsafe.Slice((*C.int32_t)(b.alloc(3*nt, 4)), 3*nt)
			for i := range faces {
				faces[i] = C.int32_t(i)
			}
			meshes[k].vertices = (*C.double)(copyTo(len(f.meshVerts[k]), 8, unsafe.Pointer(&f.meshVerts[k][0])))
			meshes[k].num_vertices = C.int64_t(3 * nt)
			meshes[k].faces = &faces[0]
			meshes[k].num_faces = C.int64_t(nt)
			meshes[k].face_materials = (*C.int32_t)(copyTo(nt, 4, unsafe.Pointer(&f.meshMats[k][0])))
			meshes[k].normals = (*C.double)(copyTo(3*nt, 8, unsafe.Pointer(&f.meshNorms[k][0])))
		}
		d.meshes, d.num_meshes = &meshes[0], C.int32_t(n)
	}
	if n := len(f.textures); n > 0 {
		tex := unsafe.Slice((*C.prt_texture)(b.alloc(n, int(unsafe.Sizeof(C.prt_texture{})))), n)
		for k, t := range f.textures {
			px := unsafe.Slice((*C.float)(b.alloc(3*len(t.Pixels), 4)), 3*len(t.Pixels))
			for i, c := range t.Pixels {
				px[3*i], px[3*i+1], px[3*i+2] = C.float(c.X), C.float(c.Y), C.float(c.Z)
			}
			tex[k] = C.prt_texture{width: C.int32_t(t.Width), height: C.int32_t(t.Height), rgb: &px[0]}
		}
		d.textures, d.num_textures = &tex[0], C.int32_t(n)
	}
	cc := s.CameraConfig
	v3(&d.camera.center, cc.Center)
	v3(&d.camera.look_at, cc.LookAt)
	v3(&d.camera.up, cc.Up)
	d.camera.width = C.int32_t(cc.Width)
	d.camera.aspect_ratio = C.double(cc.AspectRatio)
	d.camera.vfov = C.double(cc.VFov)
	d.camera.aperture = C.double(cc.Aperture)
	d.camera.focus_distance = C.double(cc.FocusDistance)
	out := &Scene{src: s}
	if err := check(C.prt_scene_create(C.int32_t(device), &d, &out.dev.h)); err != nil {
		return nil, err
	}
	// primitive ids in the library's order: mesh triangles, spheres, quads, 6 faces per box (the box
	// itself answers for its faces)
	for _, tris := range f.meshTris {
		out.prims = append(out.prims, tris...)
	}
	out.prims = append(out.prims, f.sphereSh...)
	out.prims = append(out.prims, f.quadSh...)
	for _, bx := range f.boxSh {
		for k := 0; k < 6; k++ {
			out.prims = append(out.prims, bx)
		}
	}
	for i, sh := range s.Shapes {
		if i == 0 {
			out.bbox = sh.BoundingBox()
		} else {
			out.bbox = out.bbox.Union(sh.BoundingBox())
		}
	}
	return out, nil
}

// Close frees the device scene; it fails while a film still uses it.
func (s *Scene) Close() error { return s.dev.close() }
