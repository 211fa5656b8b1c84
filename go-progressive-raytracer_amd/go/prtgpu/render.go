package prtgpu

/*
#include "prt.h"
*/
import "C"

import (
	"context"
	"errors"
	"fmt"
	"image"
	"time"
	"unsafe"

	"github.com/df07/go-progressive-raytracer/pkg/core"
	"github.com/df07/go-progressive-raytracer/pkg/integrator"
	"github.com/df07/go-progressive-raytracer/pkg/renderer"
)

// ProgressiveRaytracer mirrors renderer.ProgressiveRaytracer (pkg/renderer/progressive.go:49-96) with
// the passes rendered on the GPU: the film (per-pixel sums, luminance moments, counts and BDPT splats)
// stays on the device between passes, and images are converted there (vec3ToColor, :482-495).
type ProgressiveRaytracer struct {
	scene  *Scene
	config renderer.ProgressiveConfig
	film   *C.prt_film
	w, h   int
	tiles  []image.Rectangle
	logger core.Logger
}

// NewProgressiveRaytracer (progressive.go:62-96). The integrator selects the device estimator: a
// *integrator.PathTracingIntegrator or an *integrator.BDPTIntegrator; sampling settings come from the
// scene's SamplingConfig as in the reference. seed keys the counter RNG (renders are reproducible per seed).
func NewProgressiveRaytracer(s *Scene, config renderer.ProgressiveConfig, integ integrator.Integrator,
	logger core.Logger, seed uint64) (*ProgressiveRaytracer, error) {
	if config.TileSize <= 0 {
		return nil, errors.New("prtgpu: tile size must be positive")
	}
	var p C.prt_render_params
	switch integ.(type) {
	case *integrator.PathTracingIntegrator:
		p.integrator = C.PRT_INTEGRATOR_PATH
	case *integrator.BDPTIntegrator:
		p.integrator = C.PRT_INTEGRATOR_BDPT
	default:
		return nil, fmt.Errorf("prtgpu: integrator %T has no device implementation", integ)
	}
	sc := s.src.SamplingConfig
	p.width, p.height = C.int32_t(sc.Width), C.int32_t(sc.Height)
	p.max_depth, p.rr_min_bounces = C.int32_t(sc.MaxDepth), C.int32_t(sc.RussianRouletteMinBounces)
	p.tile_size = C.int32_t(config.TileSize)
	p.initial_samples = C.int32_t(config.InitialSamples)
	p.max_samples_per_pixel = C.int32_t(config.MaxSamplesPerPixel)
	p.max_passes = C.int32_t(config.MaxPasses)
	p.adaptive_min_samples = C.double(sc.AdaptiveMinSamples)
	p.adaptive_threshold = C.double(sc.AdaptiveThreshold)
	p.seed = C.uint64_t(seed)
	pr := &ProgressiveRaytracer{scene: s, config: config, logger: logger}
	if err := check(C.prt_film_create(s.dev.h, &p, &pr.film)); err != nil {
		return nil, err
	}
	var w, h C.int32_t
	if err := check(C.prt_film_size(pr.film, &w, &h)); err != nil {
		C.prt_film_destroy(pr.film)
		return nil, err
	}
	pr.w, pr.h = int(w), int(h)
	// NewTileGrid (progressive.go:455-480): tiles in row-major order
	ts := config.TileSize
	for y0 := 0; y0 < pr.h; y0 += ts {
		for x0 := 0; x0 < pr.w; x0 += ts {
			pr.tiles = append(pr.tiles, image.Rect(x0, y0, min(x0+ts, pr.w), min(y0+ts, pr.h)))
		}
	}
	return pr, nil
}

// Close frees the film.
func (pr *ProgressiveRaytracer) Close() {
	if pr.film != nil {
		C.prt_film_destroy(pr.film)
		pr.film = nil
	}
}

func (pr *ProgressiveRaytracer) readRGBA(r image.Rectangle) (*image.RGBA, error) {
	img := image.NewRGBA(image.Rect(0, 0, r.Dx(), r.Dy()))
	if r.Empty() {
		return img, nil
	}
	rc := C.prt_film_read_rgba8(pr.film, C.int32_t(r.Min.X), C.int32_t(r.Min.Y), C.int32_t(r.Max.X),
		C.int32_t(r.Max.Y), (*C.uint8_t)(unsafe.Pointer(&img.Pix[0])))
	return img, check(rc)
}

// RenderPass (progressive.go:126-223): renders every pixel to getSamplesForPass(pass) samples, then
// reports every tile to tileCallback (single-threaded, in tile order) and returns the image and RenderStats.
func (pr *ProgressiveRaytracer) RenderPass(pass int, tileCallback func(renderer.TileCompletionResult)) (*image.RGBA,
	renderer.RenderStats, error) {
	var st C.prt_pass_stats
	if err := check(C.prt_render_pass(pr.film, C.int32_t(pass), &st)); err != nil {
		return nil, renderer.RenderStats{}, err
	}
	img, err := pr.readRGBA(image.Rect(0, 0, pr.w, pr.h))
	if err != nil {
		return nil, renderer.RenderStats{}, err
	}
	if tileCallback != nil {
		for i, t := range pr.tiles {
			ti, err := pr.readRGBA(t)
			if err != nil {
				return nil, renderer.RenderStats{}, err
			}
			tileCallback(renderer.TileCompletionResult{
				TileX: t.Min.X / pr.config.TileSize, TileY: t.Min.Y / pr.config.TileSize, TileImage: ti,
				PassNumber: pass, TileNumber: i + 1, TotalTiles: len(pr.tiles), TotalPasses: pr.config.MaxPasses,
			})
		}
	}
	return img, renderer.RenderStats{
		TotalPixels: int(st.total_pixels), TotalSamples: int(st.total_samples),
		AverageSamples: float64(st.average_samples), MaxSamples: int(st.max_samples),
		MinSamples: int(st.min_samples), MaxSamplesUsed: int(st.max_samples_used),
	}, nil
}

// RenderProgressive (progressive.go:286-368): passes 1..MaxPasses on a goroutine, the context checked
// before each pass, results on the returned channels; the loop stops at MaxSamplesPerPixel.
func (pr *ProgressiveRaytracer) RenderProgressive(ctx context.Context, options renderer.RenderOptions) (
	<-chan renderer.PassResult, <-chan renderer.TileCompletionResult, <-chan error) {
	passChan := make(chan renderer.PassResult, 1)
	tileChan := make(chan renderer.TileCompletionResult, 100)
	errChan := make(chan error, 1)
	if !options.TileUpdates {
		close(tileChan)
	}
	go func() {
		defer close(passChan)
		if options.TileUpdates {
			defer close(tileChan)
		}
		defer close(errChan)
		for pass := 1; pass <= pr.config.MaxPasses; pass++ {
			select {
			case <-ctx.Done():
				errChan <- ctx.Err()
				return
			default:
			}
			start := time.Now()
			var cb func(renderer.TileCompletionResult)
			if options.TileUpdates {
				cb = func(r renderer.TileCompletionResult) {
					select {
					case tileChan <- r:
					case <-ctx.Done():
					default: // channel full: the tile is dropped, as the reference does
					}
				}
			}
			img, stats, err := pr.RenderPass(pass, cb)
			if err != nil {
				errChan <- err
				return
			}
			actual := int(stats.AverageSamples)
			if pr.logger != nil {
				pr.logger.Printf("Pass %d completed in %v (actual: %d samples/pixel)\n", pass, time.Since(start), actual)
			}
			res := renderer.PassResult{PassNumber: pass, Image: img, Stats: stats,
				IsLast: pass == pr.config.MaxPasses || actual >= pr.config.MaxSamplesPerPixel}
			select {
			case passChan <- res:
			case <-ctx.Done():
				return
			}
			if actual >= pr.config.MaxSamplesPerPixel {
				break
			}
		}
	}()
	return passChan, tileChan, errChan
}
