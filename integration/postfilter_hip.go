//go:build hip

// GPU post-filter runner for goav1 over libav1pf (MI355X, gfx950).
//
// Drop this file into the goav1 module root (package goav1, next to
// decoder_postfilter_bind.go: it uses the package's frame helpers) with
// libav1pf.so and include/av1pf.h under third_party/av1pf, and build with
// -tags hip.  hipPostFilterRunner implements
// internaldecoder.FrameWorkPostFilterRunner (internal/av1/decoder/work.go:153-155),
// the optional PostFilterOutput (decoder_postfilter_bind.go:252-254) and
// internaldecoder.FrameWorkPostFilterGlobalSurfacePublisher (work.go:161-163),
// so Decoder.postFilterRunner (decode_simple.go:864-869) can return it in place
// of the Go runner.  The frame pool stays in host memory: each Apply uploads
// the pre-filter reconstruction and the side data, runs the remaining stages
// in one av1pf_apply and downloads the filtered frame (and the film-grain
// display frame or the upscaled output) before returning, as the Go runner
// returns only after Output is written.  Device buffers grow monotonically
// and are reused, so steady-state frames allocate nothing.
package goav1

/*
#cgo CFLAGS: -I${SRCDIR}/third_party/av1pf/include
#cgo LDFLAGS: -L${SRCDIR}/third_party/av1pf -lav1pf -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,${SRCDIR}/third_party/av1pf
#include <stdlib.h>
#include <string.h>
#include "av1pf.h"

// The three HIP runtime calls the shim needs (hipError_t and hipMemcpyKind
// are C enums: int-sized).
extern int hipMalloc(void **ptr, size_t size);
extern int hipFree(void *ptr);
extern int hipMemcpy(void *dst, const void *src, size_t size, int kind);

// C-allocated call state: the request and every struct it points to live
// in C memory, so no Go pointer is stored in memory passed to C.
typedef struct hip_pf_call {
    av1pf_frame frame, upscaled, display;
    av1pf_request req;
    av1pf_deblock_params deblock;
    av1pf_lf_build_params build;
    av1pf_cdef_params cdef;
    av1pf_lr_params lr;
    av1pf_grain_params grain;
} hip_pf_call;
*/
import "C"

import (
	"errors"
	"unsafe"

	internaldecoder "github.com/thesyncim/goav1/internal/av1/decoder"
	internalframe "github.com/thesyncim/goav1/internal/av1/frame"
	internallfmask "github.com/thesyncim/goav1/internal/av1/lfmask"
	internalparser "github.com/thesyncim/goav1/internal/av1/parser"
	"github.com/thesyncim/goav1/internal/av1/threading"
)

const (
	hipMemcpyHostToDevice = 1
	hipMemcpyDeviceToHost = 2
)

var errHIP = errors.New("av1pf: HIP runtime call failed")

// devBuf is one device allocation, grown (never shrunk) on demand.
type devBuf struct {
	p   unsafe.Pointer
	cap int
}

func (b *devBuf) reserve(n int) error {
	if n <= b.cap {
		return nil
	}
	if b.p != nil {
		C.hipFree(b.p)
		b.p, b.cap = nil, 0
	}
	var p unsafe.Pointer
	if C.hipMalloc(&p, C.size_t(n)) != 0 {
		return errHIP
	}
	b.p, b.cap = p, n
	return nil
}

func (b *devBuf) upload(src []byte) error {
	if err := b.reserve(len(src)); err != nil || len(src) == 0 {
		return err
	}
	if C.hipMemcpy(b.p, unsafe.Pointer(&src[0]), C.size_t(len(src)), hipMemcpyHostToDevice) != 0 {
		return errHIP
	}
	return nil
}

func (b *devBuf) download(dst []byte) error {
	if len(dst) == 0 {
		return nil
	}
	if C.hipMemcpy(unsafe.Pointer(&dst[0]), b.p, C.size_t(len(dst)), hipMemcpyDeviceToHost) != 0 {
		return errHIP
	}
	return nil
}

func (b *devBuf) free() {
	if b.p != nil {
		C.hipFree(b.p)
	}
	*b = devBuf{}
}

func bytesOf[T any](s []T) []byte {
	if len(s) == 0 {
		return nil
	}
	var z T
	return unsafe.Slice((*byte)(unsafe.Pointer(&s[0])), len(s)*int(unsafe.Sizeof(z)))
}

// devFrame mirrors a host frame's three planes (whole allocations: the
// MI-aligned padding deblock and CDEF write and the borders a refreshing
// frame extends are part of the result).
type devFrame struct{ planes [3]devBuf }

func (d *devFrame) planesOf(f *Frame) [3]*internalframe.Plane {
	return [3]*internalframe.Plane{&f.Y, &f.U, &f.V}
}

// describe fills the C frame descriptor for f with d's planes (after upload
// or reserve).  frame.RequiredSize allocates alloc_height luma rows
// (frame/buffer.go:141-236); chroma planes hold alloc_height >> ss_y rows.
func (d *devFrame) describe(out *C.av1pf_frame, f *Frame, matrixIdentity bool) {
	*out = C.av1pf_frame{}
	out.width = C.int32_t(f.Format.Width)
	out.height = C.int32_t(f.Format.Height)
	out.bit_depth = C.int32_t(f.Format.BitDepth)
	if f.Format.SubsamplingX {
		out.ss_x = 1
	}
	if f.Format.SubsamplingY {
		out.ss_y = 1
	}
	if f.Format.MonoChrome {
		out.mono = 1
	}
	if matrixIdentity {
		out.matrix_identity = 1
	}
	if f.Y.Stride > 0 {
		out.alloc_height = C.int32_t(len(f.Y.Pix) / f.Y.Stride)
	}
	for i, p := range d.planesOf(f) {
		if f.Format.MonoChrome && i > 0 {
			break
		}
		out.plane[i] = d.planes[i].p
		out.stride[i] = C.int64_t(p.Stride)
	}
}

func (d *devFrame) upload(f *Frame) error {
	for i, p := range d.planesOf(f) {
		if f.Format.MonoChrome && i > 0 {
			break
		}
		if err := d.planes[i].upload(p.Pix); err != nil {
			return err
		}
	}
	return nil
}

func (d *devFrame) reserve(f *Frame) error {
	for i, p := range d.planesOf(f) {
		if f.Format.MonoChrome && i > 0 {
			break
		}
		if err := d.planes[i].reserve(len(p.Pix)); err != nil {
			return err
		}
	}
	return nil
}

func (d *devFrame) download(f *Frame) error {
	for i, p := range d.planesOf(f) {
		if f.Format.MonoChrome && i > 0 {
			break
		}
		if err := d.planes[i].download(p.Pix); err != nil {
			return err
		}
	}
	return nil
}

// deviceMirror holds the device copies of one frame's inputs and the host
// staging of the converted side data, all reused frame after frame.
type deviceMirror struct {
	output, upscaled, display devFrame
	levelCache, masks         devBuf
	lfRecords                 devBuf
	cdefIndex, noskip         devBuf
	lrUnits                   [3]devBuf

	records []C.av1pf_lf_record
	idx     []int8
	words   []uint64
	units [3][]C.av1pf_lr_unit
}

// deblock converts FrameWorkLoopFilterMasks (threading/loop_filter_masks.go:20-30):
// lfmask.FilterMask has av1pf_lf_mask's layout (1280 bytes, uint16 words) and
// the level cache is the same four bytes per MI cell.
func (m *deviceMirror) deblock(out *C.av1pf_deblock_params, masks *threading.FrameWorkLoopFilterMasks,
	lf internalparser.LoopFilterParams) error {
	var probe internallfmask.FilterMask
	if unsafe.Sizeof(probe) != uintptr(C.sizeof_av1pf_lf_mask) {
		return errors.New("av1pf: FilterMask layout differs from av1pf_lf_mask")
	}
	if err := m.masks.upload(bytesOf(masks.Masks)); err != nil {
		return err
	}
	if err := m.levelCache.upload(masks.LevelCache); err != nil {
		return err
	}
	*out = C.av1pf_deblock_params{}
	out.level_y[0], out.level_y[1] = C.uint8_t(lf.LevelY[0]), C.uint8_t(lf.LevelY[1])
	out.level_u, out.level_v = C.uint8_t(lf.LevelU), C.uint8_t(lf.LevelV)
	out.sharpness = C.uint8_t(lf.Sharpness)
	out.mi_cols, out.mi_rows = C.int32_t(masks.Cols), C.int32_t(masks.Rows)
	out.sb128_w, out.sb128_h = C.int32_t(masks.SB128W), C.int32_t(masks.SB128H)
	out.level_cache = (*C.uint8_t)(m.levelCache.p)
	out.masks = (*C.av1pf_lf_mask)(m.masks.p)
	return nil
}

// deblockFromMap builds the deblocking side data on the device from the
// loop-filter record map when no decode-time mask handle is bound
// (FrameWorkLoopFilterMasks.BuildFromMap, threading/loop_filter_masks.go:452-546):
// one av1pf_lf_record per MI cell (loop_filter_map.go:46-60) and the level
// parameters of lfmask.NewLFValueParams (lfmask/values.go:28-45).
func (m *deviceMirror) deblockFromMap(ctx *C.av1pf_ctx, stream unsafe.Pointer, out *C.av1pf_deblock_params,
	build *C.av1pf_lf_build_params, lfm *threading.FrameWorkLoopFilterMap, ev *internaldecoder.Event,
	f FrameFormat) error {
	cols, rows := int(lfm.Stride), int(lfm.Rows)
	if len(lfm.Records) < cols*rows {
		return threading.ErrInvalidBatch
	}
	m.records = m.records[:0]
	for _, rec := range lfm.Records[:cols*rows] {
		var r C.av1pf_lf_record
		r.mi_col, r.mi_row = C.uint16_t(rec.Block.MICol), C.uint16_t(rec.Block.MIRow)
		r.size = C.uint8_t(rec.Block.Size)
		r.tx_y, r.tx_uv = C.uint8_t(rec.TransformTree.Y), C.uint8_t(rec.TransformTree.UV)
		var fl C.uint8_t
		if rec.Valid {
			fl |= C.AV1PF_LF_VALID
		}
		if rec.SkipTransform {
			fl |= C.AV1PF_LF_SKIP
		}
		if rec.Intra {
			fl |= C.AV1PF_LF_INTRA
		}
		if rec.TransformTree.HasUV {
			fl |= C.AV1PF_LF_HAS_UV
		}
		r.flags = fl
		r.split[0], r.split[1] = C.uint16_t(rec.TransformTree.Split[0]), C.uint16_t(rec.TransformTree.Split[1])
		r.segment, r.ref, r.mode = C.uint8_t(rec.SegmentID), C.uint8_t(rec.RefFrame), C.uint8_t(rec.Mode)
		if ev.Delta.DeltaLFPresent { // blockLevels, loop_filter_masks.go:94-101
			if ev.Delta.DeltaLFMulti {
				for i := 0; i < 4; i++ {
					r.delta_lf[i] = C.int8_t(rec.DeltaLF[i])
				}
			} else {
				r.delta_lf[0] = C.int8_t(rec.DeltaLFFromBase)
			}
		}
		m.records = append(m.records, r)
	}
	if err := m.lfRecords.upload(bytesOf(m.records)); err != nil {
		return err
	}
	lf := ev.LoopFilter
	*build = C.av1pf_lf_build_params{}
	build.mi_cols, build.mi_rows = C.int32_t(cols), C.int32_t(rows)
	if f.SubsamplingX {
		build.ss_x = 1
	}
	if f.SubsamplingY {
		build.ss_y = 1
	}
	if f.MonoChrome {
		build.mono = 1
	}
	build.level_y[0], build.level_y[1] = C.uint8_t(lf.LevelY[0]), C.uint8_t(lf.LevelY[1])
	build.level_u, build.level_v = C.uint8_t(lf.LevelU), C.uint8_t(lf.LevelV)
	if lf.ModeRefDeltaEnabled {
		build.mode_ref_deltas = 1
	}
	if ev.Delta.DeltaLFMulti {
		build.delta_lf_multi = 1
	}
	for i := range lf.Deltas.Ref {
		build.ref_delta[i] = C.int8_t(lf.Deltas.Ref[i])
	}
	build.mode_delta[0], build.mode_delta[1] = C.int8_t(lf.Deltas.Mode[0]), C.int8_t(lf.Deltas.Mode[1])
	if ev.Segmentation.Enabled {
		build.seg_enabled = 1
		for sg := range ev.Segmentation.Data.Segments {
			d := &ev.Segmentation.Data.Segments[sg]
			build.seg_delta[sg] = [4]C.int16_t{C.int16_t(d.DeltaLFYV), C.int16_t(d.DeltaLFYH),
				C.int16_t(d.DeltaLFU), C.int16_t(d.DeltaLFV)}
		}
	}
	sbw, sbh := (cols+31)>>5, (rows+31)>>5
	if err := m.levelCache.reserve(4 * cols * rows); err != nil {
		return err
	}
	if err := m.masks.reserve(sbw * sbh * int(C.sizeof_av1pf_lf_mask)); err != nil {
		return err
	}
	if rc := C.av1pf_deblock_build(ctx, (*C.av1pf_lf_record)(m.lfRecords.p), build,
		(*C.uint8_t)(m.levelCache.p), (*C.av1pf_lf_mask)(m.masks.p), stream); rc != C.AV1PF_OK {
		return mapStatus(rc, C.GoString(C.av1pf_last_error(ctx)))
	}
	*out = C.av1pf_deblock_params{}
	out.level_y[0], out.level_y[1] = C.uint8_t(lf.LevelY[0]), C.uint8_t(lf.LevelY[1])
	out.level_u, out.level_v = C.uint8_t(lf.LevelU), C.uint8_t(lf.LevelV)
	out.sharpness = C.uint8_t(lf.Sharpness)
	out.mi_cols, out.mi_rows = C.int32_t(cols), C.int32_t(rows)
	out.sb128_w, out.sb128_h = C.int32_t(sbw), C.int32_t(sbh)
	out.level_cache = (*C.uint8_t)(m.levelCache.p)
	out.masks = (*C.av1pf_lf_mask)(m.masks.p)
	return nil
}

// cdef converts the CDEF index map (threading/cdef_index.go:13-18: Index,
// Read per 64x64 unit) to int8 indices with -1 where Read is false, and the
// CDEF skip source to one noskip word per unit: the decode-time noskip words
// of the mask handle when bound (loop_filter_noskip.go:90-101), the map's
// NoSkip words (loop_filter_noskip.go:134-139) when attached, else the
// record rule of frameWorkCDEFBlockAllSkipped (postfilter_cdef.go:1272-1300):
// an 8x8 block is skipped only when its four MI cells are valid and
// SkipTransform.
func (m *deviceMirror) cdef(out *C.av1pf_cdef_params, idx *threading.FrameWorkCDEFIndexMap,
	masks *threading.FrameWorkLoopFilterMasks, lfm *threading.FrameWorkLoopFilterMap,
	p internalparser.CDEFParams, width, height int) error {
	cols, rows := (width+63)/64, (height+63)/64
	m.idx = m.idx[:0]
	for r := 0; r < rows; r++ {
		for c := 0; c < cols; c++ {
			v := int8(-1)
			if idx != nil && r < int(idx.Rows) && c < int(idx.Stride) {
				i := r*int(idx.Stride) + c
				if i < len(idx.Read) && idx.Read[i] {
					v = int8(idx.Index[i])
				}
			}
			m.idx = append(m.idx, v)
		}
	}
	if err := m.cdefIndex.upload(bytesOf(m.idx)); err != nil {
		return err
	}
	*out = C.av1pf_cdef_params{}
	out.damping = C.uint8_t(p.Damping)
	out.strength_count = C.uint8_t(p.StrengthCount)
	for i := 0; i < 8; i++ {
		out.y_strength[i] = C.uint8_t(p.YStrength[i])
		out.uv_strength[i] = C.uint8_t(p.UVStrength[i])
	}
	out.unit_cols, out.unit_rows = C.int32_t(cols), C.int32_t(rows)
	out.index = (*C.int8_t)(m.cdefIndex.p)
	decodeTime := masks.NoSkipUnitGrid(cols, rows) // postfilter_cdef.go:595-598
	if !decodeTime && lfm == nil {
		return nil // noskip NULL: every block is filtered (postfilter_cdef.go:1277-1279)
	}
	m.words = m.words[:0]
	for r := 0; r < rows; r++ {
		for c := 0; c < cols; c++ {
			var w uint64
			switch {
			case decodeTime:
				w = masks.NoSkipUnitMask(r, c)
			case lfm.NoSkip != nil:
				w = lfm.NoSkip.UnitWord(r, c)
			default:
				for by := 0; by < 8; by++ {
					for bx := 0; bx < 8; bx++ {
						if !blockAllSkipped(lfm, r, c, by, bx) {
							w |= 1 << uint(by*8+bx)
						}
					}
				}
			}
			m.words = append(m.words, w)
		}
	}
	if err := m.noskip.upload(bytesOf(m.words)); err != nil {
		return err
	}
	out.noskip = (*C.uint64_t)(m.noskip.p)
	return nil
}

func blockAllSkipped(lfm *threading.FrameWorkLoopFilterMap, unitRow, unitCol, by, bx int) bool {
	col, row := unitCol*16+bx*2, unitRow*16+by*2
	if col+2 > int(lfm.Stride) || row+2 > int(lfm.Rows) {
		return false
	}
	for dy := 0; dy < 2; dy++ {
		for dx := 0; dx < 2; dx++ {
			rec := lfm.Records[(row+dy)*int(lfm.Stride)+col+dx]
			if !rec.Valid || !rec.SkipTransform {
				return false
			}
		}
	}
	return true
}

// lrType maps parser.RestorationType (None 0, Switchable 1, Wiener 2,
// SGRProj 3; parser/restoration.go:15-22) to AV1PF_LR_*.
func lrType(t internalparser.RestorationType) C.uint8_t {
	switch t {
	case internalparser.RestorationWiener:
		return C.AV1PF_LR_WIENER
	case internalparser.RestorationSGRProj:
		return C.AV1PF_LR_SGRPROJ
	case internalparser.RestorationSwitchable:
		return C.AV1PF_LR_SWITCHABLE
	}
	return C.AV1PF_LR_NONE
}

// restoration converts the decoded unit records (tile.RestorationUnitRecord,
// restoration_schedule.go:40-48) to row-major av1pf_lr_unit grids of the
// plan's HorzUnits x VertUnits: Wiener taps 0..2 of each direction
// (WienerFilter is symmetric with tap 3 = -2 * sum), SGR set and coded xqd.
func (m *deviceMirror) restoration(out *C.av1pf_lr_params, rb *threading.FrameWorkRestorationFrameBuffers,
	p internalparser.RestorationParams) error {
	*out = C.av1pf_lr_params{}
	for pl := 0; pl < 3; pl++ {
		out.frame_type[pl] = lrType(p.Type[pl])
		if pl == 0 {
			out.unit_size[pl] = C.uint16_t(p.UnitSizeY)
		} else {
			out.unit_size[pl] = C.uint16_t(p.UnitSizeUV)
		}
		if p.Type[pl] == internalparser.RestorationNone || rb == nil {
			continue
		}
		grid := rb.Plan.Grids[pl]
		hz, vt := int(grid.HorzUnits), int(grid.VertUnits)
		u := m.units[pl][:0]
		for i := 0; i < hz*vt; i++ {
			u = append(u, C.av1pf_lr_unit{})
		}
		for _, rec := range rb.Records[pl] {
			k := int(rec.Row)*hz + int(rec.Col)
			if k >= len(u) {
				continue
			}
			var cu C.av1pf_lr_unit
			cu._type = lrType(rec.Unit.Type)
			for t := 0; t < 3; t++ {
				cu.wiener_v[t] = C.int8_t(rec.Unit.Wiener.VFilter[t])
				cu.wiener_h[t] = C.int8_t(rec.Unit.Wiener.HFilter[t])
			}
			cu.sgr_set = C.uint8_t(rec.Unit.SGRProj.ParamsIndex)
			cu.sgr_xqd[0] = C.int8_t(rec.Unit.SGRProj.XQD[0])
			cu.sgr_xqd[1] = C.int8_t(rec.Unit.SGRProj.XQD[1])
			u[k] = cu
		}
		m.units[pl] = u
		if err := m.lrUnits[pl].upload(bytesOf(u)); err != nil {
			return err
		}
		out.units[pl] = (*C.av1pf_lr_unit)(m.lrUnits[pl].p)
	}
	return nil
}

// grain copies parser.FilmGrainParams (parser/film_grain.go:21-56) field
// for field.
func grain(out *C.av1pf_grain_params, g internalparser.FilmGrainParams) {
	*out = C.av1pf_grain_params{}
	b := func(v bool) C.uint8_t {
		if v {
			return 1
		}
		return 0
	}
	out.seed = C.uint16_t(g.Seed)
	out.num_y_points, out.num_cb_points, out.num_cr_points = C.uint8_t(g.NumYPoints), C.uint8_t(g.NumCbPoints), C.uint8_t(g.NumCrPoints)
	out.chroma_scaling_from_luma = b(g.ChromaScalingFromLuma)
	out.scaling_shift, out.ar_coeff_lag = C.uint8_t(g.ScalingShift), C.uint8_t(g.ARCoeffLag)
	out.ar_coeff_shift, out.grain_scale_shift = C.uint8_t(g.ARCoeffShift), C.uint8_t(g.GrainScaleShift)
	out.overlap, out.clip_to_restricted_range = b(g.Overlap), b(g.ClipToRestrictedRange)
	for i := range g.YPoints {
		out.y_points[i][0], out.y_points[i][1] = C.uint8_t(g.YPoints[i][0]), C.uint8_t(g.YPoints[i][1])
	}
	for i := range g.CbPoints {
		out.cb_points[i][0], out.cb_points[i][1] = C.uint8_t(g.CbPoints[i][0]), C.uint8_t(g.CbPoints[i][1])
		out.cr_points[i][0], out.cr_points[i][1] = C.uint8_t(g.CrPoints[i][0]), C.uint8_t(g.CrPoints[i][1])
	}
	for i := range g.ARCoeffsY {
		out.ar_coeffs_y[i] = C.int8_t(g.ARCoeffsY[i])
	}
	for i := range g.ARCoeffsCb {
		out.ar_coeffs_cb[i] = C.int8_t(g.ARCoeffsCb[i])
		out.ar_coeffs_cr[i] = C.int8_t(g.ARCoeffsCr[i])
	}
	out.cb_mult, out.cb_luma_mult, out.cb_offset = C.uint8_t(g.CbMult), C.uint8_t(g.CbLumaMult), C.uint16_t(g.CbOffset)
	out.cr_mult, out.cr_luma_mult, out.cr_offset = C.uint8_t(g.CrMult), C.uint8_t(g.CrLumaMult), C.uint16_t(g.CrOffset)
}

// hipPostFilterRunner is one decoder's GPU post-filter runner (the handle is
// not reentrant, as the Go runner is not: postfilter_parallel.go:143-146).
type hipPostFilterRunner struct {
	ctx    *C.av1pf_ctx
	stream unsafe.Pointer // hipStream_t, nil = the null stream
	call   *C.hip_pf_call
	dev    deviceMirror

	// super-resolution output surfaces (the external runner's pool,
	// decode_external.go:204-311) and the film-grain display frame
	outputPool *internalframe.Pool
	displayBuf []byte
	display    Frame
	source     *Frame // the frame the display frame was made from
	output     *Frame
	published  int

	// Result of the last Apply (FrameWorkPostFilterResult,
	// internal/av1/decoder/postfilter.go:309-316), as the Go runner keeps it
	// (decoder_postfilter_bind.go:237,361): the stages the call completed
	// and the loop filter's plan extent (Active, MICols, MIRows, Cells, as
	// the mask path fills them, postfilter_loopfilter_pool.go:84-93).  The
	// Go walkers' per-edge, per-unit and per-stripe counts are bookkeeping
	// of their own scalar loops and are left zero.
	Result internaldecoder.FrameWorkPostFilterResult
}

func newHIPPostFilterRunner(device int, outputPool *internalframe.Pool) (*hipPostFilterRunner, error) {
	r := &hipPostFilterRunner{outputPool: outputPool, published: -1}
	if rc := C.av1pf_create(&r.ctx, C.int(device)); rc != C.AV1PF_OK {
		return nil, errors.New("av1pf: no usable gfx950 device")
	}
	r.call = (*C.hip_pf_call)(C.calloc(1, C.size_t(C.sizeof_hip_pf_call)))
	return r, nil
}

func (r *hipPostFilterRunner) Close() {
	if r.ctx != nil {
		C.av1pf_destroy(r.ctx)
		r.ctx = nil
	}
	for _, f := range []*devFrame{&r.dev.output, &r.dev.upscaled, &r.dev.display} {
		for i := range f.planes {
			f.planes[i].free()
		}
	}
	for _, b := range []*devBuf{&r.dev.levelCache, &r.dev.masks, &r.dev.lfRecords, &r.dev.cdefIndex, &r.dev.noskip,
		&r.dev.lrUnits[0], &r.dev.lrUnits[1], &r.dev.lrUnits[2]} {
		b.free()
	}
	if r.call != nil {
		C.free(unsafe.Pointer(r.call))
		r.call = nil
	}
}

// displayFor binds the reusable display frame to format (grown, never shrunk).
func (r *hipPostFilterRunner) displayFor(format FrameFormat) (*Frame, error) {
	layout, err := internalframe.RequiredSize(format)
	if err != nil {
		return nil, err
	}
	if len(r.displayBuf) < layout.Size {
		r.displayBuf = make([]byte, layout.Size)
	}
	r.display, err = internalframe.Bind(r.displayBuf[:layout.Size], format)
	return &r.display, err
}

// Apply runs the frame's remaining post-filter stages on the device.
func (r *hipPostFilterRunner) Apply(ctx internaldecoder.FrameWorkPostFilterContext) error {
	if r == nil || r.ctx == nil || ctx.Output == nil {
		return internaldecoder.ErrInvalidFrameWorkState
	}
	r.output, r.source, r.published = nil, nil, -1
	r.Result = internaldecoder.FrameWorkPostFilterResult{}
	cl := r.call
	*cl = C.hip_pf_call{}
	remaining := ctx.RemainingPostFilters()
	identity := ctx.Event.SequenceHeader.ColorConfig.MatrixCoefficients == internalparser.MatrixCoefficientsIdentity
	out := ctx.Output
	if err := r.dev.output.upload(out); err != nil {
		return err
	}
	r.dev.output.describe(&cl.frame, out, identity)
	stages := remaining
	if !ctx.Event.FilmGrain.Apply {
		stages &^= internaldecoder.FrameWorkPostFilterFilmGrain // nothing to apply: no display frame
	}
	cl.req.stages = C.uint32_t(stages)
	if ctx.Event.FrameSize.RefreshFrameFlags != 0 {
		cl.req.flags = C.AV1PF_REQ_REFRESH // borders extended before the frame becomes a reference
	}
	if remaining.Has(internaldecoder.FrameWorkPostFilterLoopFilter) {
		// the decode-time masks when usable (loopFilterMasksUsable,
		// postfilter_loopfilter.go:499-504), else masks built on the device
		// from the shared record map, which reads across tile boundaries as
		// the edge-list apply does
		tiles := ctx.Event.TileInfo
		masks := ctx.LoopFilterMasks
		var err error
		switch {
		case masks.Valid() && ((tiles.Cols <= 1 && tiles.Rows <= 1) || masks.TileBoundariesFixed):
			err = r.dev.deblock(&cl.deblock, masks, ctx.Event.LoopFilter)
		case ctx.LoopFilterMap != nil && len(ctx.LoopFilterMap.Records) > 0:
			err = r.dev.deblockFromMap(r.ctx, r.stream, &cl.deblock, &cl.build, ctx.LoopFilterMap, &ctx.Event,
				out.Format)
		default:
			err = ErrDecoderUnsupportedPostFilter
		}
		if err != nil {
			return err
		}
		cl.req.deblock = &cl.deblock
	}
	if remaining.Has(internaldecoder.FrameWorkPostFilterCDEF) {
		if err := r.dev.cdef(&cl.cdef, ctx.CDEFIndexMap, ctx.LoopFilterMasks, ctx.LoopFilterMap, ctx.Event.CDEF, out.Format.Width,
			out.Format.Height); err != nil {
			return err
		}
		cl.req.cdef = &cl.cdef
	}
	post := out // the frame loop restoration and film grain run on
	local := -1
	if remaining.Has(internaldecoder.FrameWorkPostFilterSuperRes) {
		// caller chain (ApplyCallerPostFilters, postfilter.go:947-979): the
		// upscaled output comes from the output pool, as in the external runner
		if r.outputPool == nil {
			return internalframe.ErrInvalidPool
		}
		format, err := frameOutputFormatFromHeaders(ctx.Event.SequenceHeader, ctx.Event.FrameSize, out.Format.Align)
		if err != nil {
			return err
		}
		idx, up, err := r.outputPool.AcquireFormat(format)
		if err != nil {
			return err
		}
		local = idx
		if err := r.dev.upscaled.upload(up); err != nil {
			_ = r.outputPool.Release(local)
			return err
		}
		r.dev.upscaled.describe(&cl.upscaled, up, identity)
		cl.req.upscaled = &cl.upscaled
		post = up
	}
	release := func() {
		if local >= 0 {
			_ = r.outputPool.Release(local)
		}
	}
	if remaining.Has(internaldecoder.FrameWorkPostFilterLoopRestoration) {
		if err := r.dev.restoration(&cl.lr, ctx.RestorationFrameBuffers, ctx.Event.Restoration); err != nil {
			release()
			return err
		}
		cl.req.lr = &cl.lr
	}
	var disp *Frame
	if remaining.Has(internaldecoder.FrameWorkPostFilterFilmGrain) && ctx.Event.FilmGrain.Apply {
		var err error
		if disp, err = r.displayFor(post.Format); err != nil {
			release()
			return err
		}
		if err := r.dev.display.reserve(disp); err != nil {
			release()
			return err
		}
		r.dev.display.describe(&cl.display, disp, identity)
		grain(&cl.grain, ctx.Event.FilmGrain)
		cl.req.grain = &cl.grain
		cl.req.display = &cl.display
	}
	if rc := C.av1pf_apply(r.ctx, &cl.frame, &cl.req, r.stream); rc != C.AV1PF_OK {
		release()
		return mapStatus(rc, C.GoString(C.av1pf_last_error(r.ctx)))
	}
	if rc := C.av1pf_sync(r.ctx, r.stream); rc != C.AV1PF_OK {
		release()
		return mapStatus(rc, C.GoString(C.av1pf_last_error(r.ctx)))
	}
	if err := r.dev.output.download(out); err != nil {
		release()
		return err
	}
	r.Result.Completed = stages
	if stages.Has(internaldecoder.FrameWorkPostFilterLoopFilter) {
		lf := &r.Result.LoopFilter
		lf.Active, lf.Plan.Active = true, true
		lf.Plan.MICols, lf.Plan.MIRows = uint16(cl.deblock.mi_cols), uint16(cl.deblock.mi_rows)
		lf.Plan.Cells = int32(cl.deblock.mi_cols) * int32(cl.deblock.mi_rows)
	}
	r.output = out
	if local >= 0 {
		if err := r.dev.upscaled.download(post); err != nil {
			release()
			return err
		}
		r.output = post
		r.published = decoderExternalOutputSurfaceBase + local
	}
	if disp != nil {
		if err := r.dev.display.download(disp); err != nil {
			return err
		}
		r.output, r.source = disp, post
	}
	return nil
}

// postFilterSource maps the film-grain display frame to the reconstructed
// frame it was made from, as the Go runner's does
// (decoder_postfilter_bind.go:1078-1083), for Decoder.outputSource.
func (r *hipPostFilterRunner) postFilterSource(f *Frame) *Frame {
	if r != nil && f == &r.display && r.source != nil {
		return r.source
	}
	return f
}

// PostFilterOutput returns the display output of the last Apply: the
// film-grain frame, the upscaled frame, or the filtered frame itself.
func (r *hipPostFilterRunner) PostFilterOutput() (*Frame, bool) {
	if r == nil || r.output == nil {
		return nil, false
	}
	return r.output, true
}

// PublishedFrameWorkGlobalSurface reports the output-pool surface a
// super-resolution frame was published to (work.go:161-163).
func (r *hipPostFilterRunner) PublishedFrameWorkGlobalSurface() (int, bool) {
	if r == nil || r.published < 0 {
		return -1, false
	}
	return r.published, true
}

func mapStatus(rc C.int, msg string) error {
	switch rc {
	case C.AV1PF_ERR_UNSUPPORTED:
		return ErrDecoderUnsupportedPostFilter
	case C.AV1PF_ERR_SHORT_BUFFER:
		return internalframe.ErrShortBuffer
	case C.AV1PF_ERR_INVALID_ARG:
		return errors.New("av1pf: invalid request: " + msg)
	}
	return errors.New("av1pf: " + msg)
}

var (
	_ internaldecoder.FrameWorkPostFilterRunner                 = (*hipPostFilterRunner)(nil)
	_ decoderFrameWorkPostFilterOutputProvider                  = (*hipPostFilterRunner)(nil)
	_ internaldecoder.FrameWorkPostFilterGlobalSurfacePublisher = (*hipPostFilterRunner)(nil)
)
