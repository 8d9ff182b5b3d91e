//! `gpu-rocm` routing of the reference's Reed-Solomon data path to `librsgpu.so`.
//!
//! The reference crate adds this directory as `src/gpu/` (`mod gpu;` in
//! `src/lib.rs` behind `#[cfg(feature = "gpu-rocm")]`) with `build.rs` beside its
//! own.  `ReedSolomon` methods whose shards live in HIP device memory forward to
//! `GpuCodec`; the stream methods forward to `stream_io`.  Device-resident shards
//! are passed as `rsgpu_cslice` / `rsgpu_slice` / `rsgpu_slot` (pointer + length
//! of device memory); host-resident shards use the `*_host` calls, which stage
//! through pinned memory inside the library.
//!
//! Every call validates all of its arguments before any device work is enqueued,
//! in the reference's order, and returns the reference's `crate::Error`.  Device
//! failures (status 100 and up) have no `crate::Error` variant and panic with the
//! library's description.
#![allow(dead_code)]

pub mod ffi;
pub mod stream_io;

use core::ffi::{c_void, CStr};
use std::io::{Read, Write};

use ffi::*;

/// Status -> `crate::Error`: codes 1..=16 follow `src/errors.rs:4-21` in order.
pub(crate) fn map_status(rc: i32) -> Result<(), crate::Error> {
    use crate::Error::*;
    const E: [crate::Error; 16] = [
        TooFewShards, TooManyShards, TooFewDataShards, TooManyDataShards, TooFewParityShards,
        TooManyParityShards, TooFewBufferShards, TooManyBufferShards, IncorrectShardSize,
        TooFewShardsPresent, EmptyShard, InvalidShardFlags, InvalidIndex, InvalidCustomMatrix,
        UnsupportedCodecFamily, UnsupportedLeopardPrototype,
    ];
    match rc {
        0 => Ok(()),
        1..=16 => Err(E[(rc - 1) as usize]),
        _ => {
            let name = unsafe { CStr::from_ptr(rsgpu_status_name(rc)) }.to_string_lossy();
            let hip = unsafe { CStr::from_ptr(rsgpu_last_hip_error()) }.to_string_lossy();
            panic!("rsgpu device error {rc} ({name}): {hip}");
        }
    }
}

/// The HIP stream work is enqueued on; null = the default stream.
pub type Stream = *mut c_void;

/// A codec handle.  The library's codec is safe to share across threads: its
/// decode caches sit behind mutexes, like `inversion_cache.rs:38-48`.
pub struct GpuCodec {
    raw: *mut rsgpu_codec,
}

unsafe impl Send for GpuCodec {}
unsafe impl Sync for GpuCodec {}

impl Drop for GpuCodec {
    fn drop(&mut self) {
        unsafe { rsgpu_codec_free(self.raw) }
    }
}

fn flags(v: &[bool]) -> Vec<u8> {
    v.iter().map(|&b| b as u8).collect()
}

impl GpuCodec {
    /// `ReedSolomon::new` / `with_options`: `opts` from `rsgpu_options_default`
    /// with the fields of the reference's `CodecOptions` set.
    pub fn new(data_shards: usize, parity_shards: usize, opts: &rsgpu_options) -> Result<Self, crate::Error> {
        let mut raw = std::ptr::null_mut();
        map_status(unsafe { rsgpu_codec_new(data_shards, parity_shards, opts, &mut raw) })?;
        Ok(Self { raw })
    }

    pub fn with_defaults(data_shards: usize, parity_shards: usize) -> Result<Self, crate::Error> {
        let mut o = rsgpu_options::default();
        unsafe { rsgpu_options_default(&mut o) };
        Self::new(data_shards, parity_shards, &o)
    }

    /// `ReedSolomon::with_custom_matrix`: `rows` is n_rows x row_len, row-major.
    pub fn with_custom_matrix(
        data_shards: usize,
        parity_shards: usize,
        rows: &[u8],
        n_rows: usize,
        row_len: usize,
        opts: &rsgpu_options,
    ) -> Result<Self, crate::Error> {
        let mut raw = std::ptr::null_mut();
        map_status(unsafe {
            rsgpu_codec_with_custom_matrix(data_shards, parity_shards, rows.as_ptr(), n_rows, row_len, opts, &mut raw)
        })?;
        Ok(Self { raw })
    }

    pub fn raw(&self) -> *const rsgpu_codec {
        self.raw
    }

    // ---- encode.rs
    pub fn encode(&self, shards: &[rsgpu_slice], stream: Stream) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_encode(self.raw, shards.as_ptr(), shards.len(), stream) })
    }

    pub fn encode_sep(&self, data: &[rsgpu_cslice], parity: &[rsgpu_slice], stream: Stream) -> Result<(), crate::Error> {
        map_status(unsafe {
            rsgpu_encode_sep(self.raw, data.as_ptr(), data.len(), parity.as_ptr(), parity.len(), stream)
        })
    }

    pub fn encode_single(&self, i_data: usize, shards: &[rsgpu_slice], stream: Stream) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_encode_single(self.raw, i_data, shards.as_ptr(), shards.len(), stream) })
    }

    pub fn encode_single_sep(
        &self,
        i_data: usize,
        single: rsgpu_cslice,
        parity: &[rsgpu_slice],
        stream: Stream,
    ) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_encode_single_sep(self.raw, i_data, single, parity.as_ptr(), parity.len(), stream) })
    }

    /// `update` (encode.rs:1048): `None` new data = an unchanged shard.
    pub fn update(
        &self,
        old_data: &[rsgpu_cslice],
        new_data: &[Option<rsgpu_cslice>],
        parity: &[rsgpu_slice],
        stream: Stream,
    ) -> Result<(), crate::Error> {
        let new: Vec<rsgpu_cslice> = new_data
            .iter()
            .map(|s| s.unwrap_or(rsgpu_cslice { ptr: std::ptr::null(), len: 0 }))
            .collect();
        map_status(unsafe {
            rsgpu_update(self.raw, old_data.as_ptr(), old_data.len(), new.as_ptr(), new.len(), parity.as_ptr(),
                         parity.len(), stream)
        })
    }

    // ---- verify.rs (waits for the stream: the answer is a host bool)
    pub fn verify(&self, shards: &[rsgpu_cslice], stream: Stream) -> Result<bool, crate::Error> {
        let mut ok = 0i32;
        map_status(unsafe { rsgpu_verify(self.raw, shards.as_ptr(), shards.len(), &mut ok, stream) })?;
        Ok(ok != 0)
    }

    pub fn verify_with_buffer(
        &self,
        shards: &[rsgpu_cslice],
        buffer: &[rsgpu_slice],
        stream: Stream,
    ) -> Result<bool, crate::Error> {
        let mut ok = 0i32;
        map_status(unsafe {
            rsgpu_verify_with_buffer(self.raw, shards.as_ptr(), shards.len(), buffer.as_ptr(), buffer.len(), &mut ok,
                                     stream)
        })?;
        Ok(ok != 0)
    }

    // ---- reconstruct.rs: `present` of each slot plays the Option / ShardSlot role
    pub fn reconstruct(&self, shards: &mut [rsgpu_slot], stream: Stream) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_reconstruct(self.raw, shards.as_mut_ptr(), shards.len(), stream) })
    }

    pub fn reconstruct_data(&self, shards: &mut [rsgpu_slot], stream: Stream) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_reconstruct_data(self.raw, shards.as_mut_ptr(), shards.len(), stream) })
    }

    pub fn reconstruct_some(&self, shards: &mut [rsgpu_slot], required: &[bool], stream: Stream) -> Result<(), crate::Error> {
        let req = flags(required);
        map_status(unsafe {
            rsgpu_reconstruct_some(self.raw, shards.as_mut_ptr(), shards.len(), req.as_ptr(), req.len(), stream)
        })
    }

    /// `prepare_reconstruct_workspace` (reconstruct.rs:539-700).
    pub fn prepare_reconstruct_workspace(&self, layout: &[rsgpu_slot], stream: Stream) -> Result<Workspace, crate::Error> {
        let mut raw = std::ptr::null_mut();
        map_status(unsafe { rsgpu_prepare_reconstruct_workspace(self.raw, layout.as_ptr(), layout.len(), stream, &mut raw) })?;
        Ok(Workspace { raw })
    }

    pub fn prepare_reconstruct_data_workspace(&self, layout: &[rsgpu_slot], stream: Stream) -> Result<Workspace, crate::Error> {
        let mut raw = std::ptr::null_mut();
        map_status(unsafe {
            rsgpu_prepare_reconstruct_data_workspace(self.raw, layout.as_ptr(), layout.len(), stream, &mut raw)
        })?;
        Ok(Workspace { raw })
    }

    pub fn prepare_reconstruct_some_workspace(
        &self,
        layout: &[rsgpu_slot],
        required: &[bool],
        stream: Stream,
    ) -> Result<Workspace, crate::Error> {
        let req = flags(required);
        let mut raw = std::ptr::null_mut();
        map_status(unsafe {
            rsgpu_prepare_reconstruct_some_workspace(self.raw, layout.as_ptr(), layout.len(), req.as_ptr(), req.len(),
                                                     stream, &mut raw)
        })?;
        Ok(Workspace { raw })
    }

    /// `reconstruct_with_workspace` and `reconstruct_with_workspace_par`.
    pub fn reconstruct_with_workspace(
        &self,
        shards: &mut [rsgpu_slot],
        workspace: &Workspace,
        stream: Stream,
    ) -> Result<(), crate::Error> {
        map_status(unsafe {
            rsgpu_reconstruct_with_workspace(self.raw, shards.as_mut_ptr(), shards.len(), workspace.raw, stream)
        })
    }

    /// `decode_idx` (galois_8/policy.rs:543-704): `None` destinations and inputs
    /// are null slices; `expect_input: None` is the merge mode.
    pub fn decode_idx(
        &self,
        dst: &[Option<rsgpu_slice>],
        expect_input: Option<&[bool]>,
        input: &[Option<rsgpu_cslice>],
        stream: Stream,
    ) -> Result<(), crate::Error> {
        let d: Vec<rsgpu_slice> =
            dst.iter().map(|s| s.unwrap_or(rsgpu_slice { ptr: std::ptr::null_mut(), len: 0 })).collect();
        let i: Vec<rsgpu_cslice> =
            input.iter().map(|s| s.unwrap_or(rsgpu_cslice { ptr: std::ptr::null(), len: 0 })).collect();
        let e = expect_input.map(flags);
        let (ep, en) = e.as_ref().map_or((std::ptr::null(), 0), |e| (e.as_ptr(), e.len()));
        map_status(unsafe { rsgpu_decode_idx(self.raw, d.as_ptr(), d.len(), ep, en, i.as_ptr(), i.len(), stream) })
    }

    /// No reference counterpart: reconstruct without any allocation, for callers
    /// that cannot allocate (HIP graph capture, tight serving loops).  Calls that
    /// overlap on one GPU share its scratch and race.
    pub fn reconstruct_noalloc(&self, shards: &mut [rsgpu_slot], data_only: bool, stream: Stream) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_reconstruct_noalloc(self.raw, shards.as_mut_ptr(), shards.len(), data_only as i32, stream) })
    }

    // ---- codec.rs split / join
    pub fn split_len(&self, data_len: usize) -> Result<usize, crate::Error> {
        let mut n = 0usize;
        map_status(unsafe { rsgpu_split_len(self.raw, data_len, &mut n) })?;
        Ok(n)
    }

    // ---- host-resident shards (the reference's usual case), PCIe-staged
    pub fn encode_host(&self, shards: &mut [&mut [u8]]) -> Result<(), crate::Error> {
        let len = shards.first().map_or(0, |s| s.len());
        let ptrs: Vec<*mut u8> = shards.iter_mut().map(|s| s.as_mut_ptr()).collect();
        map_status(unsafe { rsgpu_encode_host(self.raw, ptrs.as_ptr(), ptrs.len(), len) })
    }

    pub fn verify_host(&self, shards: &[&[u8]]) -> Result<bool, crate::Error> {
        let len = shards.first().map_or(0, |s| s.len());
        let ptrs: Vec<*const u8> = shards.iter().map(|s| s.as_ptr()).collect();
        let mut ok = 0i32;
        map_status(unsafe { rsgpu_verify_host(self.raw, ptrs.as_ptr(), ptrs.len(), len, &mut ok) })?;
        Ok(ok != 0)
    }

    /// `present[i]` is updated to 1 for every rebuilt shard.
    pub fn reconstruct_host(&self, shards: &mut [&mut [u8]], present: &mut [u8], data_only: bool) -> Result<(), crate::Error> {
        let len = shards.first().map_or(0, |s| s.len());
        let ptrs: Vec<*mut u8> = shards.iter_mut().map(|s| s.as_mut_ptr()).collect();
        map_status(unsafe {
            rsgpu_reconstruct_host(self.raw, ptrs.as_ptr(), present.as_mut_ptr(), ptrs.len(), len, data_only as i32)
        })
    }

    // ---- stream.rs
    pub fn encode_stream<R: Read + Send, W: Write + Send>(
        &self,
        data: &mut [R],
        parity: &mut [W],
        opts: &crate::stream::StreamOptions,
    ) -> Result<(), crate::stream::StreamError> {
        stream_io::encode_stream(self.raw, data, parity, opts)
    }

    pub fn verify_stream<R: Read + Send>(
        &self,
        shards: &mut [R],
        opts: &crate::stream::StreamOptions,
    ) -> Result<bool, crate::stream::StreamError> {
        stream_io::verify_stream(self.raw, shards, opts)
    }

    pub fn reconstruct_stream_from<R: Read + Send, W: Write + Send>(
        &self,
        readers: &mut [Option<R>],
        writers: &mut [Option<W>],
        opts: &crate::stream::StreamOptions,
    ) -> Result<(), crate::stream::StreamError> {
        stream_io::reconstruct_stream_from(self.raw, readers, writers, opts)
    }

    pub fn reconstruct_stream(
        &self,
        shards: &mut [std::io::Cursor<Vec<u8>>],
        opts: &crate::stream::StreamOptions,
    ) -> Result<(), crate::stream::StreamError> {
        stream_io::reconstruct_stream(self.raw, shards, opts)
    }
}

/// A prepared reconstruction plan (`ReconstructWorkspace`).
pub struct Workspace {
    raw: *mut rsgpu_reconstruct_workspace,
}

unsafe impl Send for Workspace {}
unsafe impl Sync for Workspace {}

impl Drop for Workspace {
    fn drop(&mut self) {
        unsafe { rsgpu_reconstruct_workspace_free(self.raw) }
    }
}

impl Workspace {
    /// (`data_targets()`, `parity_targets()`).
    pub fn targets(&self) -> (Vec<usize>, Vec<usize>) {
        let (mut nd, mut np) = (0usize, 0usize);
        let _ = unsafe { rsgpu_reconstruct_workspace_targets(self.raw, std::ptr::null_mut(), 0, &mut nd, &mut np) };
        let mut t = vec![0usize; nd + np];
        let _ = unsafe { rsgpu_reconstruct_workspace_targets(self.raw, t.as_mut_ptr(), t.len(), &mut nd, &mut np) };
        let parity = t.split_off(nd);
        (t, parity)
    }
}

/// Leopard plug points (leopard.rs:277-317): the same arguments as
/// `leopard_gf8_encode` / `leopard_gf16_encode` on device-resident slices.
pub fn leopard_encode(
    bits: i32,
    data_shards: usize,
    parity_shards: usize,
    data: &[rsgpu_cslice],
    parity: &[rsgpu_slice],
    stream: Stream,
) -> Result<(), crate::Error> {
    map_status(unsafe {
        rsgpu_leopard_encode(bits, data_shards, parity_shards, data.as_ptr(), data.len(), parity.as_ptr(), parity.len(),
                             stream)
    })
}

/// `reconstruct_with_tables`: an empty output slice is a shard not requested.
pub fn leopard_reconstruct(
    bits: i32,
    data_shards: usize,
    parity_shards: usize,
    present: &[bool],
    inputs: &[rsgpu_cslice],
    outputs: &[rsgpu_slice],
    stream: Stream,
) -> Result<(), crate::Error> {
    let p = flags(present);
    map_status(unsafe {
        rsgpu_leopard_reconstruct(bits, data_shards, parity_shards, p.as_ptr(), inputs.as_ptr(), outputs.as_ptr(), stream)
    })
}

/// The request server (continuous batching under an HBM budget).  Outputs of a
/// completed request live in the server's arena until `release`.
pub struct Server {
    raw: *mut rsgpu_server,
}

unsafe impl Send for Server {}

impl Drop for Server {
    fn drop(&mut self) {
        unsafe { rsgpu_server_free(self.raw) }
    }
}

impl Server {
    pub fn new(config: &rsgpu_server_config) -> Result<Self, crate::Error> {
        let mut raw = std::ptr::null_mut();
        map_status(unsafe { rsgpu_server_new(config, &mut raw) })?;
        Ok(Self { raw })
    }

    pub fn submit(&mut self, request: &rsgpu_request) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_server_submit(self.raw, request) })
    }

    /// Admits, launches and retires what it can; returns the batches in flight.
    pub fn pump(&mut self) -> Result<u32, crate::Error> {
        let mut n = 0u32;
        map_status(unsafe { rsgpu_server_pump(self.raw, &mut n) })?;
        Ok(n)
    }

    pub fn poll(&mut self, max: usize) -> Result<Vec<rsgpu_completion>, crate::Error> {
        let mut out = vec![rsgpu_completion::default(); max];
        let mut n = 0usize;
        map_status(unsafe { rsgpu_server_poll(self.raw, out.as_mut_ptr(), max, &mut n) })?;
        out.truncate(n);
        Ok(out)
    }

    /// Device pointers of a completed request's outputs (n of them).
    pub fn outputs(&self, tag: u64, n: usize) -> Result<Vec<*mut u8>, crate::Error> {
        let mut out = vec![std::ptr::null_mut(); n];
        map_status(unsafe { rsgpu_server_outputs(self.raw, tag, out.as_mut_ptr(), n) })?;
        Ok(out)
    }

    /// Returns a completed request's outputs to the arena.
    pub fn release(&mut self, tag: u64) -> Result<(), crate::Error> {
        map_status(unsafe { rsgpu_server_release(self.raw, tag) })
    }

    pub fn replay(&mut self, requests: &[rsgpu_request], keep_outputs: bool) -> Result<Vec<rsgpu_completion>, crate::Error> {
        let mut out = vec![rsgpu_completion::default(); requests.len()];
        map_status(unsafe {
            rsgpu_server_replay(self.raw, requests.as_ptr(), requests.len(), keep_outputs as i32, out.as_mut_ptr())
        })?;
        Ok(out)
    }

    pub fn stats(&self) -> Result<rsgpu_server_counters, crate::Error> {
        let mut s = rsgpu_server_counters::default();
        map_status(unsafe { rsgpu_server_stats(self.raw, &mut s) })?;
        Ok(s)
    }
}
