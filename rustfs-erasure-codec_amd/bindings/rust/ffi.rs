//! Raw bindings of `librsgpu.so` (`include/rsgpu.h`), one item per declaration of
//! the header.  The reference crate adds this file as `src/gpu/ffi.rs` behind a
//! `gpu-rocm` feature; `build.rs` beside it builds and links the library.
//!
//! Status codes 1..=16 are the reference's `crate::Error` variants in
//! declaration order (`src/errors.rs:4-21`); 100.. are device and I/O failures.
#![allow(non_camel_case_types)]
#![allow(dead_code)]

use core::ffi::{c_char, c_int, c_void};

pub const RSGPU_OK: c_int = 0;
pub const RSGPU_ERR_HIP: c_int = 100;
pub const RSGPU_ERR_INVALID_ARGUMENT: c_int = 101;
pub const RSGPU_ERR_NO_DEVICE: c_int = 102;
pub const RSGPU_ERR_UNSUPPORTED: c_int = 103;
pub const RSGPU_ERR_OUT_OF_MEMORY: c_int = 104;
pub const RSGPU_ERR_IO: c_int = 105;

pub const RSGPU_FIELD_GF8: i32 = 0;
pub const RSGPU_FIELD_GF16: i32 = 1;
pub const RSGPU_FAMILY_CLASSIC: i32 = 0;
pub const RSGPU_FAMILY_LEOPARD_GF8: i32 = 1;
pub const RSGPU_FAMILY_LEOPARD_GF16: i32 = 2;
pub const RSGPU_MATRIX_VANDERMONDE: i32 = 0;
pub const RSGPU_MATRIX_CAUCHY: i32 = 1;
pub const RSGPU_MATRIX_JERASURE_LIKE: i32 = 2;
pub const RSGPU_MATRIX_CUSTOM: i32 = 3;
pub const RSGPU_LEOPARD_DISABLED: i32 = 0;
pub const RSGPU_LEOPARD_AS_NEEDED: i32 = 1;
pub const RSGPU_LEOPARD_PREFER_GF16: i32 = 2;
pub const RSGPU_LEOPARD_PREFER_LEOPARD: i32 = 3;

pub const RSGPU_STREAM_OK: i32 = 0;
pub const RSGPU_STREAM_READ: i32 = 1;
pub const RSGPU_STREAM_WRITE: i32 = 2;
pub const RSGPU_STREAM_CODEC: i32 = 3;
pub const RSGPU_STREAM_IO_AUTO: i32 = 0;
pub const RSGPU_STREAM_IO_SERIAL: i32 = 1;
pub const RSGPU_STREAM_IO_PARALLEL: i32 = 2;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rsgpu_options {
    pub field: i32,
    pub codec_family: i32,
    pub matrix_mode: i32,
    pub leopard_mode: i32,
    pub fast_one_parity: i32,
    pub inversion_cache: i32,
    pub inversion_cache_capacity: u64,
    pub inversion_cache_max_bytes: u64,
}

#[repr(C)]
pub struct rsgpu_codec {
    _p: [u8; 0],
}
#[repr(C)]
pub struct rsgpu_reconstruct_workspace {
    _p: [u8; 0],
}
#[repr(C)]
pub struct rsgpu_batch {
    _p: [u8; 0],
}
#[repr(C)]
pub struct rsgpu_server {
    _p: [u8; 0],
}
#[repr(C)]
pub struct rsgpu_mlp {
    _p: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_cslice {
    pub ptr: *const u8,
    pub len: usize,
}
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_slice {
    pub ptr: *mut u8,
    pub len: usize,
}
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_slot {
    pub ptr: *mut u8,
    pub len: usize,
    pub present: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_stream_options {
    pub block_size: usize,
    pub io_mode: i32,
    pub reserved: i32,
}
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rsgpu_stream_error {
    pub kind: i32,
    pub shard_index: i32,
    pub io_error: i64,
}
pub type rsgpu_read_cb =
    unsafe extern "C" fn(user: *mut c_void, shard: usize, buf: *mut u8, len: usize) -> i64;
pub type rsgpu_write_cb =
    unsafe extern "C" fn(user: *mut c_void, shard: usize, buf: *const u8, len: usize) -> i64;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_mem_stream {
    pub data: *mut u8,
    pub len: usize,
    pub pos: usize,
}
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_mem_streams {
    pub readers: *mut rsgpu_mem_stream,
    pub writers: *mut rsgpu_mem_stream,
}
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_vec {
    pub data: *mut u8,
    pub len: usize,
    pub cap: usize,
    pub pos: usize,
}
pub type rsgpu_grow_cb =
    unsafe extern "C" fn(user: *mut c_void, shard: usize, vec: *mut rsgpu_vec, need: usize) -> c_int;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rsgpu_server_config {
    pub hbm_budget: u64,
    pub arena_bytes: u64,
    pub max_inflight_batches: i32,
    pub max_batch_ops: i32,
}
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rsgpu_request {
    pub op: i32,
    pub reserved: i32,
    pub codec: *const rsgpu_codec,
    pub shard_len: usize,
    pub shards: *const *const u8,
    pub erased: *const u8,
    pub arrival_ns: u64,
    pub tag: u64,
}
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rsgpu_completion {
    pub tag: u64,
    pub status: i32,
    pub verify_ok: i32,
    pub arrival_ns: u64,
    pub admit_ns: u64,
    pub launch_ns: u64,
    pub done_ns: u64,
    pub footprint: u64,
}
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct rsgpu_server_counters {
    pub completed: u64,
    pub batches: u64,
    pub launched_ops: u64,
    pub max_inflight_batches: u64,
    pub peak_footprint: u64,
    pub footprint: u64,
}

#[link(name = "rsgpu")]
extern "C" {
    // ---- codecs (ReedSolomon::new / with_options / with_custom_matrix)
    pub fn rsgpu_options_default(opts: *mut rsgpu_options);
    pub fn rsgpu_codec_new(
        data_shards: usize,
        parity_shards: usize,
        opts: *const rsgpu_options,
        out: *mut *mut rsgpu_codec,
    ) -> c_int;
    pub fn rsgpu_codec_with_custom_matrix(
        data_shards: usize,
        parity_shards: usize,
        rows: *const u8,
        n_rows: usize,
        row_len: usize,
        opts: *const rsgpu_options,
        out: *mut *mut rsgpu_codec,
    ) -> c_int;
    pub fn rsgpu_codec_free(codec: *mut rsgpu_codec);
    pub fn rsgpu_codec_info(
        codec: *const rsgpu_codec,
        data_shards: *mut usize,
        parity_shards: *mut usize,
        field: *mut i32,
        codec_family: *mut i32,
    ) -> c_int;
    pub fn rsgpu_codec_matrix(codec: *const rsgpu_codec, out: *mut u8, out_len: usize) -> c_int;
    pub fn rsgpu_codec_format_id(
        codec: *const rsgpu_codec,
        out: *mut u8,
        cap: usize,
        len: *mut usize,
    ) -> c_int;

    // ---- split / join (codec.rs:476, 507)
    pub fn rsgpu_split_len(codec: *const rsgpu_codec, data_len: usize, shard_len: *mut usize) -> c_int;
    pub fn rsgpu_split(
        codec: *const rsgpu_codec,
        data: *const u8,
        data_len: usize,
        shards: *const rsgpu_slice,
        n_shards: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_join(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_cslice,
        n_shards: usize,
        out_len: usize,
        out: *mut u8,
        written: *mut usize,
        stream: *mut c_void,
    ) -> c_int;

    // ---- encode (encode.rs)
    pub fn rsgpu_encode(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_slice,
        n_shards: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_encode_sep(
        codec: *const rsgpu_codec,
        data: *const rsgpu_cslice,
        n_data: usize,
        parity: *const rsgpu_slice,
        n_parity: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_encode_sep_gf2mfma(
        codec: *const rsgpu_codec,
        data: *const rsgpu_cslice,
        n_data: usize,
        parity: *const rsgpu_slice,
        n_parity: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_encode_single(
        codec: *const rsgpu_codec,
        i_data: usize,
        shards: *const rsgpu_slice,
        n_shards: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_encode_single_sep(
        codec: *const rsgpu_codec,
        i_data: usize,
        single: rsgpu_cslice,
        parity: *const rsgpu_slice,
        n_parity: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_update(
        codec: *const rsgpu_codec,
        old_data: *const rsgpu_cslice,
        n_old: usize,
        new_data: *const rsgpu_cslice,
        n_new: usize,
        parity: *const rsgpu_slice,
        n_parity: usize,
        stream: *mut c_void,
    ) -> c_int;

    // ---- verify (verify.rs)
    pub fn rsgpu_verify(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_cslice,
        n_shards: usize,
        ok: *mut i32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_verify_async(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_cslice,
        n_shards: usize,
        device_flag: *mut i32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_verify_with_buffer(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_cslice,
        n_shards: usize,
        buffer: *const rsgpu_slice,
        n_buffer: usize,
        ok: *mut i32,
        stream: *mut c_void,
    ) -> c_int;

    // ---- reconstruct (reconstruct.rs)
    pub fn rsgpu_reconstruct(
        codec: *const rsgpu_codec,
        shards: *mut rsgpu_slot,
        n_shards: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_reconstruct_data(
        codec: *const rsgpu_codec,
        shards: *mut rsgpu_slot,
        n_shards: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_reconstruct_some(
        codec: *const rsgpu_codec,
        shards: *mut rsgpu_slot,
        n_shards: usize,
        required: *const u8,
        n_required: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_prepare_reconstruct_workspace(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_slot,
        n_shards: usize,
        stream: *mut c_void,
        out: *mut *mut rsgpu_reconstruct_workspace,
    ) -> c_int;
    pub fn rsgpu_prepare_reconstruct_data_workspace(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_slot,
        n_shards: usize,
        stream: *mut c_void,
        out: *mut *mut rsgpu_reconstruct_workspace,
    ) -> c_int;
    pub fn rsgpu_prepare_reconstruct_some_workspace(
        codec: *const rsgpu_codec,
        shards: *const rsgpu_slot,
        n_shards: usize,
        required: *const u8,
        n_required: usize,
        stream: *mut c_void,
        out: *mut *mut rsgpu_reconstruct_workspace,
    ) -> c_int;
    pub fn rsgpu_reconstruct_with_workspace(
        codec: *const rsgpu_codec,
        shards: *mut rsgpu_slot,
        n_shards: usize,
        workspace: *const rsgpu_reconstruct_workspace,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_reconstruct_workspace_targets(
        workspace: *const rsgpu_reconstruct_workspace,
        targets: *mut usize,
        cap: usize,
        n_data: *mut usize,
        n_parity: *mut usize,
    ) -> c_int;
    pub fn rsgpu_reconstruct_workspace_free(workspace: *mut rsgpu_reconstruct_workspace);
    pub fn rsgpu_decode_idx(
        codec: *const rsgpu_codec,
        dst: *const rsgpu_slice,
        n_dst: usize,
        expect_input: *const u8,
        n_expect: usize,
        input: *const rsgpu_cslice,
        n_input: usize,
        stream: *mut c_void,
    ) -> c_int;
    /// No reference counterpart: for callers that cannot allocate.
    pub fn rsgpu_reconstruct_noalloc(
        codec: *const rsgpu_codec,
        shards: *mut rsgpu_slot,
        n_shards: usize,
        data_only: i32,
        stream: *mut c_void,
    ) -> c_int;

    // ---- Leopard plug points (leopard.rs:277-317)
    pub fn rsgpu_leopard_encode(
        bits: i32,
        data_shards: usize,
        parity_shards: usize,
        data: *const rsgpu_cslice,
        n_data: usize,
        parity: *const rsgpu_slice,
        n_parity: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_leopard_reconstruct(
        bits: i32,
        data_shards: usize,
        parity_shards: usize,
        present: *const u8,
        inputs: *const rsgpu_cslice,
        outputs: *const rsgpu_slice,
        stream: *mut c_void,
    ) -> c_int;

    // ---- block streams (stream.rs:677-1031)
    pub fn rsgpu_stream_options_default(options: *mut rsgpu_stream_options);
    pub fn rsgpu_encode_stream(
        codec: *const rsgpu_codec,
        n_data: usize,
        n_parity: usize,
        read: rsgpu_read_cb,
        write: rsgpu_write_cb,
        user: *mut c_void,
        options: *const rsgpu_stream_options,
        err: *mut rsgpu_stream_error,
    ) -> c_int;
    pub fn rsgpu_verify_stream(
        codec: *const rsgpu_codec,
        n_shards: usize,
        read: rsgpu_read_cb,
        user: *mut c_void,
        options: *const rsgpu_stream_options,
        ok: *mut i32,
        err: *mut rsgpu_stream_error,
    ) -> c_int;
    pub fn rsgpu_reconstruct_stream_from(
        codec: *const rsgpu_codec,
        n_readers: usize,
        n_writers: usize,
        has_reader: *const u8,
        has_writer: *const u8,
        read: rsgpu_read_cb,
        write: rsgpu_write_cb,
        user: *mut c_void,
        options: *const rsgpu_stream_options,
        err: *mut rsgpu_stream_error,
    ) -> c_int;
    pub fn rsgpu_mem_stream_read(streams: *mut c_void, shard: usize, buf: *mut u8, len: usize) -> i64;
    pub fn rsgpu_mem_stream_write(streams: *mut c_void, shard: usize, buf: *const u8, len: usize) -> i64;
    pub fn rsgpu_reconstruct_stream(
        codec: *const rsgpu_codec,
        shards: *mut rsgpu_vec,
        n_shards: usize,
        grow: Option<rsgpu_grow_cb>,
        user: *mut c_void,
        options: *const rsgpu_stream_options,
        err: *mut rsgpu_stream_error,
    ) -> c_int;

    // ---- shards in host memory: chunked pinned pipeline
    pub fn rsgpu_encode_host(
        codec: *const rsgpu_codec,
        shards: *const *mut u8,
        n_shards: usize,
        shard_len: usize,
    ) -> c_int;
    pub fn rsgpu_verify_host(
        codec: *const rsgpu_codec,
        shards: *const *const u8,
        n_shards: usize,
        shard_len: usize,
        ok: *mut i32,
    ) -> c_int;
    pub fn rsgpu_reconstruct_host(
        codec: *const rsgpu_codec,
        shards: *const *mut u8,
        present: *mut u8,
        n_shards: usize,
        shard_len: usize,
        data_only: i32,
    ) -> c_int;

    // ---- SelectedEncodeKernel::execute (galois_8/codegen.rs:48)
    pub fn rsgpu_apply_rows(
        field: i32,
        rows: *const u8,
        r: usize,
        k: usize,
        input: *const rsgpu_cslice,
        out: *const rsgpu_slice,
        stream: *mut c_void,
    ) -> c_int;

    // ---- batches (encode_sep_par / execute_classic_plan_parallel over many stripes)
    pub fn rsgpu_batch_new(out: *mut *mut rsgpu_batch) -> c_int;
    pub fn rsgpu_batch_free(batch: *mut rsgpu_batch);
    pub fn rsgpu_batch_clear(batch: *mut rsgpu_batch) -> c_int;
    pub fn rsgpu_batch_add_encode(
        batch: *mut rsgpu_batch,
        codec: *const rsgpu_codec,
        shards: *const rsgpu_slice,
        n_shards: usize,
    ) -> c_int;
    pub fn rsgpu_batch_add_verify(
        batch: *mut rsgpu_batch,
        codec: *const rsgpu_codec,
        shards: *const rsgpu_cslice,
        n_shards: usize,
        device_flag: *mut i32,
    ) -> c_int;
    pub fn rsgpu_batch_add_reconstruct(
        batch: *mut rsgpu_batch,
        codec: *const rsgpu_codec,
        shards: *mut rsgpu_slot,
        n_shards: usize,
        data_only: i32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_batch_set_tune(
        batch: *mut rsgpu_batch,
        inputs_per_group: i32,
        vectors_per_lane: i32,
        blocks_per_cu: i32,
    ) -> c_int;
    pub fn rsgpu_batch_autotune(
        batch: *mut rsgpu_batch,
        stream: *mut c_void,
        reps: i32,
        inputs_per_group: *mut i32,
        vectors_per_lane: *mut i32,
        blocks_per_cu: *mut i32,
        ms: *mut f32,
    ) -> c_int;
    pub fn rsgpu_batch_run(batch: *mut rsgpu_batch, stream: *mut c_void) -> c_int;
    pub fn rsgpu_batch_graph_launch(batch: *mut rsgpu_batch, stream: *mut c_void) -> c_int;
    pub fn rsgpu_batch_stats(
        batch: *const rsgpu_batch,
        n_ops: *mut u64,
        bytes_moved: *mut u64,
        data_bytes: *mut u64,
        tiles: *mut u64,
    ) -> c_int;

    // ---- kernel-variant dispatch
    pub fn rsgpu_variant_count() -> c_int;
    pub fn rsgpu_dispatch_dims(n_ops: *mut i32, n_buckets: *mut i32, bucket0: *mut i32) -> c_int;
    pub fn rsgpu_variant_info(
        variant: i32,
        kind: *mut i32,
        lanes: *mut i32,
        vectors_per_lane: *mut i32,
        rows_per_group: *mut i32,
        leopard_tile_words: *mut i32,
    ) -> c_int;
    pub fn rsgpu_variant_valid(codec: *const rsgpu_codec, op: i32, erasures: usize, variant: i32) -> c_int;
    pub fn rsgpu_codec_set_variant(codec: *mut rsgpu_codec, variant: i32) -> c_int;
    pub fn rsgpu_codec_set_variant_table(codec: *mut rsgpu_codec, table: *const i16, n_erasures: usize) -> c_int;
    pub fn rsgpu_codec_variant_for(
        codec: *const rsgpu_codec,
        op: i32,
        erasures: usize,
        shard_len: usize,
        variant: *mut i32,
    ) -> c_int;

    // ---- request server
    pub fn rsgpu_server_new(config: *const rsgpu_server_config, out: *mut *mut rsgpu_server) -> c_int;
    pub fn rsgpu_server_free(server: *mut rsgpu_server);
    pub fn rsgpu_server_submit(server: *mut rsgpu_server, request: *const rsgpu_request) -> c_int;
    pub fn rsgpu_server_pump(server: *mut rsgpu_server, batches_in_flight: *mut u32) -> c_int;
    pub fn rsgpu_server_poll(
        server: *mut rsgpu_server,
        out: *mut rsgpu_completion,
        max: usize,
        n: *mut usize,
    ) -> c_int;
    pub fn rsgpu_server_outputs(
        server: *const rsgpu_server,
        tag: u64,
        outputs: *mut *mut u8,
        n: usize,
    ) -> c_int;
    pub fn rsgpu_server_release(server: *mut rsgpu_server, tag: u64) -> c_int;
    pub fn rsgpu_server_replay(
        server: *mut rsgpu_server,
        requests: *const rsgpu_request,
        n: usize,
        keep_outputs: i32,
        out: *mut rsgpu_completion,
    ) -> c_int;
    pub fn rsgpu_server_stats(server: *const rsgpu_server, stats: *mut rsgpu_server_counters) -> c_int;

    // ---- the learned dispatcher's network
    pub fn rsgpu_mlp_new(n_in: i32, n_out: i32, seed: u64, out: *mut *mut rsgpu_mlp) -> c_int;
    pub fn rsgpu_mlp_free(mlp: *mut rsgpu_mlp);
    pub fn rsgpu_mlp_dims(
        mlp: *const rsgpu_mlp,
        n_in: *mut i32,
        n_hidden: *mut i32,
        n_out: *mut i32,
        n_params: *mut usize,
    ) -> c_int;
    pub fn rsgpu_mlp_get_params(mlp: *const rsgpu_mlp, params: *mut f32, n_params: usize) -> c_int;
    pub fn rsgpu_mlp_set_params(mlp: *mut rsgpu_mlp, params: *const f32, n_params: usize) -> c_int;
    pub fn rsgpu_mlp_forward(
        mlp: *mut rsgpu_mlp,
        x: *const f32,
        batch: usize,
        y: *mut f32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_mlp_loss_grad(
        mlp: *mut rsgpu_mlp,
        x: *const f32,
        t: *const f32,
        w: *const f32,
        batch: usize,
        loss: *mut f32,
        grad: *mut f32,
        n_params: usize,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_mlp_train(
        mlp: *mut rsgpu_mlp,
        x: *const f32,
        t: *const f32,
        w: *const f32,
        batch: usize,
        steps: i32,
        lr: f32,
        loss: *mut f32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn rsgpu_codec_install_dispatch(
        codec: *mut rsgpu_codec,
        net: *mut rsgpu_mlp,
        envelope: *const i64,
        stream: *mut c_void,
    ) -> c_int;

    // ---- diagnostics
    pub fn rsgpu_cache_stats(
        codec: *const rsgpu_codec,
        requests: *mut u64,
        hits: *mut u64,
        misses: *mut u64,
        entries: *mut u64,
    ) -> c_int;
    pub fn rsgpu_status_name(status: c_int) -> *const c_char;
    pub fn rsgpu_init(device: c_int) -> c_int;
    pub fn rsgpu_last_hip_error() -> *const c_char;
}
