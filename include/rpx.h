/*
 * rpx — the C ABI of the MI355X tracer (librpx.so), a drop-in for the reference renderer's
 * render call.
 *
 * The first group mirrors the reference's hand-written wasm C ABI one for one (same names,
 * argument meaning, status codes and ownership rules); paths are inside the reference's
 * engine/crates/rt-wasm/src:
 *   rt_version            ffi.rs:12-15
 *   rt_features           ffi.rs:16-25
 *   rt_alloc              ffi.rs:27-33   (len zeroed bytes of host memory; never null)
 *   rt_free               ffi.rs:35-48   (a buffer of rt_alloc with its length; null ignored)
 *   rt_scene_load         ffi.rs:52-75   (RSB v1 bytes, copied; 0 on error)
 *   rt_scene_update       ffi.rs:77-88   (keeps the meshes when the bytes carry no MESH)
 *   rt_scene_error_len    ffi.rs:90-93
 *   rt_scene_error_copy   ffi.rs:95-111
 *   rt_scene_free         ffi.rs:113-124
 *   rt_render_tile        ffi.rs:118-158 (one sample of a tile: w*h*3 linear RGB floats)
 *   rt_render_tile_rgba   ffi.rs:160-196 (w*h*4 floats, RGB premultiplied, alpha)
 *   rt_render_block_rgba  ffi.rs:196-234 (a block of an adaptive round: one sample of the
 *                         rectangle, as rt_render_tile_rgba; `iteration` picks the photon map
 *                         of a scene with caustics; scenes with edge supersampling are refused at
 *                         load)
 *   rt_scene_caustics     ffi.rs:420-425 (caustic photons per sample: 0 without caustics, none
 *                         asked for, nothing that focuses light or no light that sends photons;
 *                         UINT32_MAX for a bad handle). A scene with caustics traces the photon
 *                         map of each sample index on the GPU before rendering it and keeps the
 *                         last 16 maps (rt-wasm/src/lib.rs:58,103-121), so tiles of one sample
 *                         rendered one by one trace its photons once.
 *   rt_adaptive_decide    ffi.rs:237-296 (the adaptive decision of a width x height image from its
 *                         blocks' sample counts, RGBA sums (4 floats per pixel) and squared-
 *                         luminance sums: one byte per block (1: in the next round) and six words:
 *                         the next round's samples, active blocks, blocks, then the f32 bits of the
 *                         threshold, the largest error and the remaining work; status 2 for an
 *                         empty image). The per-pixel errors, their 3 x 3 smoothing and the
 *                         per-block maxima run on the GPU.
 *   rt_render_tile_passes ffi.rs:303-338 (the exact passes of `mask` of a tile, from one ray
 *                         through each pixel's centre: normal 2 (3 f32), depth 4, part id 8 (u32),
 *                         material id 16 (u32), circle of confusion 32, planes one after another
 *                         in bit order, little-endian words; pixels outside the image background)
 *   rt_render_tile_albedo ffi.rs:340-380 (w*h*3 floats: the base colour of what sample
 *                         `sample_index`'s camera ray meets first; 0 for the background)
 *   rt_render_tile_alpha  ffi.rs:382-420 (w*h floats: 1 where that ray meets an instance)
 *   rt_denoise            ffi.rs:600-636 (the denoiser, full filter, on a band of `rows` rows from
 *                         `origin_y` of a width x height image of `samples` samples: in_ptr holds
 *                         width*rows*14 little-endian words, the planes sums (RGBA), lum2, albedo
 *                         (3), normal (3), depth, part id, material id; writes the denoised RGBA
 *                         sums of rows y0..y1; status 5 when the sizes, the band or its halo do
 *                         not fit)
 *   rt_denoise_with       ffi.rs:638-687 (preset 0 full, 1 viewport; strength 0..1 blends the
 *                         noisy rows with the denoised ones, 0 runs no filter)
 *   rt_denoise_halo       ffi.rs:689-692 (rows a band holds around its rows: 72 full, 24 viewport)
 * Status codes (rt-wasm/src/lib.rs:36-51): 0 OK, 1 bad handle, 2 empty or oversized tile
 * (more than 2048 x 2048 pixels), 3 null pointer, 4 bad scene update, 5 bad denoise input; codes
 * from 16 on are the
 * GPU's (16 no device or device memory exhausted, 17 kernel launch or execution failure).
 * Output buffers belong to the caller and need no alignment.
 *
 * The second group (rpx_*) has no counterpart in the reference: it is the GPU-native form of the
 * hosts' progressive loop (rt-cli/src/lib.rs:266-366), with the film resident in HBM. A film
 * accumulates the samples of each call in sample-index order per pixel (image.rs:81-94), so any
 * split of [0, n) into consecutive calls gives the same film. Calls on one handle are
 * serialised; different handles run on their own streams (two per handle: the path pool's, and
 * one its shadow rays are traced on beside the closest hit, joined within each call).
 */
#ifndef RPX_H
#define RPX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPX_OK 0u
#define RPX_BAD_HANDLE 1u
#define RPX_BAD_TILE 2u
#define RPX_NULL_POINTER 3u
#define RPX_BAD_SCENE 4u
#define RPX_BAD_DENOISE 5u
#define RPX_GPU_MEMORY 16u
#define RPX_GPU_LAUNCH 17u

uint32_t rt_version(void);
uint32_t rt_features(void);
uint8_t* rt_alloc(uint32_t len);
void rt_free(uint8_t* ptr, uint32_t len);
uint32_t rt_scene_load(const uint8_t* ptr, uint32_t len);
uint32_t rt_scene_update(uint32_t handle, const uint8_t* ptr, uint32_t len);
uint32_t rt_scene_error_len(void);
void rt_scene_error_copy(uint8_t* ptr);
uint32_t rt_scene_free(uint32_t handle);
uint32_t rt_scene_caustics(uint32_t handle);
uint32_t rt_render_tile(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                        float* out);
uint32_t rt_render_tile_rgba(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                             uint32_t sample_index, float* out);
uint32_t rt_render_block_rgba(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                              uint32_t sample_index, uint32_t iteration, float* out);
uint32_t rt_adaptive_decide(uint32_t width, uint32_t height, float ev, float target, uint32_t cap,
                            const uint8_t* counts_ptr, const uint8_t* sums_ptr, const uint8_t* lum2_ptr,
                            uint8_t* active_ptr, uint8_t* decision_ptr);
uint32_t rt_render_tile_passes(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t mask,
                               uint8_t* out_ptr);
uint32_t rt_render_tile_albedo(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                               float* out_ptr);
uint32_t rt_render_tile_alpha(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                              float* out_ptr);
uint32_t rt_denoise(uint32_t width, uint32_t height, uint32_t origin_y, uint32_t rows, uint32_t samples, uint32_t y0,
                    uint32_t y1, const uint8_t* in_ptr, float* out_ptr);
uint32_t rt_denoise_with(uint32_t preset, float strength, uint32_t width, uint32_t height, uint32_t origin_y,
                         uint32_t rows, uint32_t samples, uint32_t y0, uint32_t y1, const uint8_t* in_ptr,
                         float* out_ptr);
uint32_t rt_denoise_halo(uint32_t preset);

/* Renders samples [first, first + count) of every pixel into the handle's resident film. */
uint32_t rpx_render_samples(uint32_t handle, uint32_t first, uint32_t count);
/* Clears the resident film. */
uint32_t rpx_film_reset(uint32_t handle);
/* Copies the film's sums to host buffers (3*W*H, W*H, W*H, W*H floats; any may be null). */
uint32_t rpx_film_read(uint32_t handle, float* rgb, float* weight, float* alpha, float* lum2);
/* The film's device pointers (rgb, weight, alpha, lum2), valid until the handle is freed. */
uint32_t rpx_film_device(uint32_t handle, void** ptrs4);
/* Nearest and any hit of n rays (8 floats each: origin, direction, t_min, t_max) in host
 * memory: instance, mesh triangle (0xffffffff on a miss), t (inf on a miss), any (0/1). */
uint32_t rpx_trace_rays(uint32_t handle, const float* rays, uint32_t n, uint32_t* inst, uint32_t* tri, float* t,
                        uint8_t* any);
/* The same queries traced one wavefront per ray (the traversal the drain gives the last paths of a
 * scene with media; no counterpart in the reference): 16 BVH entries per round over a shared stack
 * of stack_entries (1..1024) before a ray falls back to one lane. Results equal rpx_trace_rays'. */
uint32_t rpx_trace_rays_wave(uint32_t handle, const float* rays, uint32_t n, uint32_t stack_entries, uint32_t* inst,
                             uint32_t* tri, float* t, uint8_t* any);
/* width, height, max_samples, seed, max_depth, triangles, BVH nodes, instances. */
uint32_t rpx_scene_info(uint32_t handle, uint64_t* out8);
/* Since the handle was loaded: rays cast (camera + extension + shadow, as the reference counts
 * them), of them closest-hit and shadow rays, and samples rendered. */
uint32_t rpx_scene_stats(uint32_t handle, uint64_t* out4);
/* Pool options of a handle's renders: path slots (0 keeps the current), closest-hit and shadow
 * launch shapes (index 0..7; 0xffffffff keeps), iterations per completion check (0 keeps), most
 * path samples per chunk (0 keeps; at most 2^27), flags: bit 0 one sample pass at a time (the
 * pool drains after each pass), bit 1 each group of iterations launched as a captured HIP graph,
 * bit 2 no drain, bit 3 adaptive rounds render only their own samples (default: a block's samples
 * up to the cap are rendered ahead in one run once it is a candidate, and kept); tail: once no new path is left, at most this many live paths are finished in
 * one drain launch, each lane taking its path to its end (0 keeps; default 2^16). Films do not
 * depend on any of them. */
uint32_t rpx_set_options(uint32_t handle, uint32_t pool, uint32_t shape, uint32_t shadow_shape, uint32_t group,
                         uint32_t chunk, uint32_t flags, uint32_t tail);
/* The launch tuner: renders samples [0, samples) of the image (default 32; the film is untouched)
 * with every closest-hit launch shape (8) at pool sizes 2^21 .. 2^26, then every shadow launch
 * shape at the best of those, then the closest-hit trip budgets 0 and 48 at the best of those,
 * keeps the fastest as the handle's options, and reports: pool, closest-hit shape, shadow shape,
 * best ms of the shape stages, then the 48 + 8 shape candidates' ms (rpx_trip_budget reads the
 * chosen budget). */
uint32_t rpx_autotune(uint32_t handle, uint32_t samples, float* out60);
/* The closest-hit trip budget (no counterpart in the reference, whose CPU walks are not split): a
 * closest-hit walk that has made `trips` loop trips (node visits and leaves) is set aside, with its
 * stack, for a second launch that takes such walks to their ends, so that the first launch's
 * waves are not held by a few long walks. 0: one launch. The hits, and so the films, do not depend
 * on it (the closest hit is the least (t, instance, triangle), whatever the order of visits).
 * Default 48; the resident-lane launch shapes (5 .. 7) always run unsplit. rpx_tune_trip_budget
 * renders samples [0, samples) (default 32) at the handle's launch options with budgets 0 and 48,
 * keeps the faster and reports it, then both candidates' ms. */
uint32_t rpx_set_trip_budget(uint32_t handle, uint32_t trips);
uint32_t rpx_trip_budget(uint32_t handle, uint32_t* out);
uint32_t rpx_tune_trip_budget(uint32_t handle, uint32_t samples, float* out3);
/* Records per-stage GPU timing of the following renders (events on the handle's stream). */
uint32_t rpx_set_profiling(uint32_t handle, uint32_t on);
/* Timing of the last render: total ms, then ms of the closest-hit, shade, shadow and regen
 * stages (summed; with profiling on), their launch counts, pool iterations and chunks. The
 * shadow kernel runs beside the closest hit on a second stream: the closest-hit stage spans
 * both traversals, the shadow figure is the shadow kernel's own time within it. */
uint32_t rpx_last_timing(uint32_t handle, float* out12);
/* Kernel times of the last render with profiling on (ms summed over its launches): the
 * closest-hit kernel alone, its launches, the shadow kernel (on the second stream), its launches,
 * the traversal phase (closest-hit kernel until the shadow kernel beside it has joined), the
 * media walks, the drain, the shade, the photon maps of a scene with caustics (tracing and
 * building), and the photon maps the scene has traced since it was loaded. */
uint32_t rpx_last_kernel_ms(uint32_t handle, float* out10);
/* The GPU-generated energy-compensation tables of the current device (1024, 1024, 17408,
 * 17408 u16 and 32, 32 f32). */
uint32_t rpx_energy_tables(uint16_t* albedo, uint16_t* schlick, uint16_t* clear_in, uint16_t* clear_out,
                           float* albedo_avg, float* schlick_avg);
/* Scene generators (this library's workloads), written as RSB v1 bytes into buf (cap bytes);
 * returns the length, or 0 when cap is too small (call with cap 0 to size the buffer).
 * kind 0: the reference's example scene (example.rs:113-230); 1: icosphere on a floor
 * (256x256); 2: eight displaced level-7 icospheres (1024x1024); 4: the inverse-material fit's scene
 * (200,000 triangles, 512x512, floor material 8 takes the base-colour texture); 5: a glass bottle
 * holding a scattering medium, lit by a sphere light (200,000 triangles, 1920x1080); 6: that bottle
 * of clear glass under a glossy disc light, 4,194,304 caustic photons per sample. */
uint64_t rpx_write_scene(uint32_t kind, uint32_t seed, uint8_t* buf, uint64_t cap);
/* Job `job` of the seeded mixed render queue (this library's batched workload), as RSB v1 bytes
 * (same buffer protocol): a square image of 256..1024 px, 10k..500k triangles, 16/32/64 spp. */
uint64_t rpx_write_queue_job(uint32_t seed, uint32_t job, uint8_t* buf, uint64_t cap);
/* Differentiable rendering (no counterpart in the reference). Five parameters per material, in
 * this order: GGX roughness, base colour r, g, b, IOR. rpx_material_params returns the number of
 * materials and writes their parameters (up to cap floats); rpx_set_material_params replaces all
 * of them (n = number of materials). rpx_set_texture gives one material an RGB f32 base-colour
 * texture (w x h, row-major, bilinear, repeating), mapped planar in world x, y:
 * u = (x - origin_x) * scale_x, v = (y - origin_y) * scale_y; material 0xffffffff removes it.
 * rpx_adjoint: gradients of a loss of the image (the mean of samples [first, first + count))
 * given d loss / d image (3*W*H floats), by path replay backpropagation: d loss / d parameter
 * (5 per material) and, with a texture, d loss / d texel (3*w*h floats; may be null). */
uint32_t rpx_material_params(uint32_t handle, float* out, uint32_t cap);
uint32_t rpx_set_material_params(uint32_t handle, const float* params, uint32_t n);
uint32_t rpx_set_texture(uint32_t handle, uint32_t material, uint32_t w, uint32_t h, const float* rgb, float origin_x,
                         float origin_y, float scale_x, float scale_y);
uint32_t rpx_adjoint(uint32_t handle, uint32_t first, uint32_t count, const float* d_image, float* grad_materials,
                     float* grad_texture);
/* Render queues: jobs (scene handle, samples [first, first + count) of its whole image) rendered
 * together by one persistent path pool, each into its own resident film. Films equal those of
 * rpx_render_samples on the same scene and samples. */
uint32_t rpx_queue_create(void); /* 0 on error */
uint32_t rpx_queue_add(uint32_t queue, uint32_t scene, uint32_t first, uint32_t count, uint32_t* job);
uint32_t rpx_queue_run(uint32_t queue); /* renders every job once more (films accumulate) */
uint32_t rpx_queue_reset(uint32_t queue); /* clears every job's film */
uint32_t rpx_queue_film_read(uint32_t queue, uint32_t job, float* rgb, float* weight, float* alpha, float* lum2);
uint32_t rpx_queue_set_options(uint32_t queue, uint32_t pool, uint32_t shape, uint32_t shadow_shape,
                               uint32_t group, uint32_t chunk, uint32_t flags, uint32_t tail);
/* The tuner over the queue's jobs' first samples (up to `paths` path samples; 0: 2^25). */
uint32_t rpx_queue_autotune(uint32_t queue, uint32_t paths, float* out60);
/* The closest-hit trip budget the queue tuner chose for the queue (as rpx_set_trip_budget; default 48). */
uint32_t rpx_queue_trip_budget(uint32_t queue, uint32_t* out);
uint32_t rpx_queue_set_profiling(uint32_t queue, uint32_t on);
uint32_t rpx_queue_last_timing(uint32_t queue, float* out12);
uint32_t rpx_queue_stats(uint32_t queue, uint64_t* out4);
uint32_t rpx_queue_free(uint32_t queue);
/* Adaptive sampling in rounds (the reference hosts' loop over rt_render_block_rgba and
 * rt_adaptive_decide, rt-core/src/adaptive.rs:319-554, run with the film resident in HBM): clears
 * the handle's film, renders a warm-up round of every 8 x 8 block, then rounds of the blocks the
 * decisions choose, each block's samples in index order, until no block is chosen; `cap` bounds
 * every block's samples (0: the scene's max_samples), the scene's noise_threshold is the target and
 * its exposure the floor of the mean. out8: decisions, rounds, path samples, fewest and most
 * samples of a block, the f32 bits of the last decision's largest error and remaining work, and the
 * GPU time in ns. rpx_adaptive_counts copies the last adaptive render's samples per block (row by
 * row; status 2 when `cap` words do not hold them). */
uint32_t rpx_render_adaptive(uint32_t handle, uint32_t cap, uint64_t* out8);
uint32_t rpx_adaptive_counts(uint32_t handle, uint32_t* counts, uint32_t cap);
/* The last adaptive render's work (no counterpart in the reference): path samples rendered, of
 * them rendered ahead of their rounds, runs that rendered ahead, and samples rendered ahead that no
 * round used. */
uint32_t rpx_adaptive_work(uint32_t handle, uint64_t* out4);
/* Passes and the denoiser with everything resident in HBM. rpx_render_passes writes the exact
 * passes of every pixel (centre rays) into the handle's pass film and adds the albedo and coverage
 * of samples [first, first + count) to its sums in sample order (rt-core/src/render.rs:500-640);
 * rpx_passes_reset clears the sums; rpx_passes_read copies the resolved planes of `mask` (bits as
 * rt_render_tile_passes, plus albedo 1 and alpha 64) one after another in bit order.
 * rpx_denoise_film denoises the resident film (its sums of `samples` samples; 0: the samples
 * rendered since the last reset) with the resident passes, preset and strength as rt_denoise_with,
 * into W*H*4 RGBA sums (out may be null: the result stays on the device); rpx_denoise_timing: total
 * ms, then setup, firefly, variance, a-trous and final ms, levels, 0. rpx_denoise_counted is
 * rt_denoise_with over 15 words per pixel: the 14 planes and each pixel's own sample count (f32),
 * the input of an adaptive image (rt-wasm/src/lib.rs:369-475, DENOISE_WORDS_COUNTED). */
uint32_t rpx_render_passes(uint32_t handle, uint32_t first, uint32_t count);
uint32_t rpx_passes_reset(uint32_t handle);
uint32_t rpx_passes_read(uint32_t handle, uint32_t mask, uint8_t* out);
uint32_t rpx_denoise_film(uint32_t handle, uint32_t preset, float strength, uint32_t samples, float* out);
uint32_t rpx_denoise_timing(uint32_t handle, float* out8);
uint32_t rpx_denoise_counted(uint32_t preset, float strength, uint32_t width, uint32_t height, uint32_t origin_y,
                             uint32_t rows, uint32_t samples, uint32_t y0, uint32_t y1, const uint8_t* in_ptr,
                             float* out_ptr);
/* Allocation-free variant of rt_render_tile (no counterpart in the reference): for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, output (w*h*3 linear RGB floats of one sample, host memory)
 * and film bits as rt_render_tile, for tiles of at most RPX_NOALLOC_MAX_PIXELS pixels (status 2
 * beyond). The path state, results and staging live in one 16 MiB region of static device memory
 * per GPU, held until the process ends; the call allocates nothing. Every call on a GPU shares
 * that GPU's region, so calls that overlap on one GPU (from different threads or handles) race:
 * serialise them. */
#define RPX_NOALLOC_MAX_PIXELS 16384u
uint32_t rpx_render_tile_noalloc(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                                 uint32_t sample_index, float* out);
/* Number of HIP devices visible (0 without a GPU). */
uint32_t rpx_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
