//! Block streams (`encode_stream`, `verify_stream`, `reconstruct_stream_from`,
//! src/core/stream.rs:677-1031) over the library's callback entry points.
//!
//! The caller's `Read` / `Write` streams stay in Rust: the library calls back into
//! `read_cb` / `write_cb` below for each block, with a pointer to `Streams` as
//! `user`.  The library stages blocks through pinned memory and overlaps these
//! callbacks with the device copies and kernels.
//!
//! Errors come back as the reference's `StreamError`: the shard index the library
//! reports, with the `io::Error` the stream returned (kept per shard while the
//! call runs) or the codec's `crate::Error`.
#![allow(dead_code)]

use core::ffi::c_void;
use std::io::{self, Read, Write};
use std::panic::{catch_unwind, AssertUnwindSafe};
use std::sync::Mutex;

use super::ffi::*;
use super::map_status;
use crate::stream::{StreamError, StreamErrorKind, StreamIoMode, StreamOptions};

/// What the callbacks reach through `user`.  In parallel I/O mode the library
/// calls the callbacks of different shards from several threads at once and
/// never two calls for one shard at a time, so each `*mut` below is used by one
/// thread at a time; the error slots are behind mutexes.
struct Streams<R, W> {
    readers: Vec<*mut R>, // null: no reader at this index
    writers: Vec<*mut W>, // null: no writer at this index
    read_err: Vec<Mutex<Option<io::Error>>>,
    write_err: Vec<Mutex<Option<io::Error>>>,
}

// SAFETY: see the struct comment; R and W are Send, as the reference requires.
unsafe impl<R: Send, W: Send> Sync for Streams<R, W> {}

impl<R, W> Streams<R, W> {
    fn new(readers: Vec<*mut R>, writers: Vec<*mut W>) -> Self {
        let read_err = (0..readers.len()).map(|_| Mutex::new(None)).collect();
        let write_err = (0..writers.len()).map(|_| Mutex::new(None)).collect();
        Self { readers, writers, read_err, write_err }
    }

    fn user(&self) -> *mut c_void {
        self as *const Self as *mut c_void
    }
}

fn stash(slot: &Mutex<Option<io::Error>>, e: io::Error) {
    if let Ok(mut g) = slot.lock() {
        *g = Some(e);
    }
}

/// read_one_stream's inner step (stream.rs:340-352): `Interrupted` is retried,
/// any other error is kept for the StreamError and reported as -1.
unsafe extern "C" fn read_cb<R: Read, W>(user: *mut c_void, shard: usize, buf: *mut u8, len: usize) -> i64 {
    let s = &*(user as *const Streams<R, W>);
    let r = catch_unwind(AssertUnwindSafe(|| {
        let reader = &mut *s.readers[shard];
        let out = std::slice::from_raw_parts_mut(buf, len);
        loop {
            match reader.read(out) {
                Ok(n) => return Ok(n),
                Err(e) if e.kind() == io::ErrorKind::Interrupted => continue,
                Err(e) => return Err(e),
            }
        }
    }));
    match r {
        Ok(Ok(n)) => n as i64,
        Ok(Err(e)) => {
            stash(&s.read_err[shard], e);
            -1
        }
        Err(_) => {
            stash(&s.read_err[shard], io::Error::new(io::ErrorKind::Other, "reader panicked"));
            -1
        }
    }
}

/// write_block's inner step (stream.rs:473-493): a write of 0 bytes is reported as
/// 0, which the library turns into a write error (`WriteZero` below).
unsafe extern "C" fn write_cb<R, W: Write>(user: *mut c_void, shard: usize, buf: *const u8, len: usize) -> i64 {
    let s = &*(user as *const Streams<R, W>);
    let r = catch_unwind(AssertUnwindSafe(|| {
        let writer = &mut *s.writers[shard];
        let data = std::slice::from_raw_parts(buf, len);
        loop {
            match writer.write(data) {
                Ok(n) => return Ok(n),
                Err(e) if e.kind() == io::ErrorKind::Interrupted => continue,
                Err(e) => return Err(e),
            }
        }
    }));
    match r {
        Ok(Ok(n)) => n as i64,
        Ok(Err(e)) => {
            stash(&s.write_err[shard], e);
            -1
        }
        Err(_) => {
            stash(&s.write_err[shard], io::Error::new(io::ErrorKind::Other, "writer panicked"));
            -1
        }
    }
}

fn options(o: &StreamOptions) -> rsgpu_stream_options {
    rsgpu_stream_options {
        block_size: o.block_size,
        io_mode: match o.io_mode {
            StreamIoMode::Auto => RSGPU_STREAM_IO_AUTO,
            StreamIoMode::Serial => RSGPU_STREAM_IO_SERIAL,
            StreamIoMode::Parallel => RSGPU_STREAM_IO_PARALLEL,
        },
        reserved: 0,
    }
}

/// The StreamError for a call's status.  `writer_base` maps the reported index of
/// a write error back to the writer's own index (encode reports parity j as k + j).
fn result<R, W>(rc: i32, err: &rsgpu_stream_error, s: &Streams<R, W>, writer_base: usize) -> Result<(), StreamError> {
    let index = err.shard_index as usize;
    let take = |slot: Option<&Mutex<Option<io::Error>>>| slot.and_then(|m| m.lock().ok()?.take());
    match rc {
        RSGPU_OK => Ok(()),
        RSGPU_ERR_IO if err.kind == RSGPU_STREAM_READ => {
            let e = take(s.read_err.get(index))
                .unwrap_or_else(|| io::Error::new(io::ErrorKind::Other, format!("read returned {}", err.io_error)));
            Err(StreamError { shard_index: index, kind: StreamErrorKind::Read(e) })
        }
        RSGPU_ERR_IO => {
            let e = take(s.write_err.get(index.wrapping_sub(writer_base)))
                .unwrap_or_else(|| io::Error::from(io::ErrorKind::WriteZero));
            Err(StreamError { shard_index: index, kind: StreamErrorKind::Write(e) })
        }
        _ => match map_status(rc) {
            Err(e) => Err(StreamError { shard_index: index, kind: StreamErrorKind::Codec(e) }),
            Ok(()) => Ok(()),
        },
    }
}

/// `ReedSolomon::encode_stream` (stream.rs:677-752).
pub fn encode_stream<R: Read + Send, W: Write + Send>(
    codec: *const rsgpu_codec,
    data: &mut [R],
    parity: &mut [W],
    opts: &StreamOptions,
) -> Result<(), StreamError> {
    let s: Streams<R, W> = Streams::new(
        data.iter_mut().map(|r| r as *mut R).collect(),
        parity.iter_mut().map(|w| w as *mut W).collect(),
    );
    let o = options(opts);
    let mut err = rsgpu_stream_error::default();
    let rc = unsafe {
        rsgpu_encode_stream(codec, data.len(), parity.len(), read_cb::<R, W>, write_cb::<R, W>, s.user(), &o,
                            &mut err)
    };
    result(rc, &err, &s, data.len())
}

/// `ReedSolomon::verify_stream` (stream.rs:760-812).
pub fn verify_stream<R: Read + Send>(
    codec: *const rsgpu_codec,
    shards: &mut [R],
    opts: &StreamOptions,
) -> Result<bool, StreamError> {
    let s: Streams<R, io::Sink> = Streams::new(shards.iter_mut().map(|r| r as *mut R).collect(), Vec::new());
    let o = options(opts);
    let mut err = rsgpu_stream_error::default();
    let mut ok = 0i32;
    let rc = unsafe {
        rsgpu_verify_stream(codec, shards.len(), read_cb::<R, io::Sink>, s.user(), &o, &mut ok, &mut err)
    };
    result(rc, &err, &s, 0).map(|()| ok != 0)
}

/// `ReedSolomon::reconstruct_stream_from` (stream.rs:948-1031): `Some` readers are
/// the present shards, `Some` writers receive the rebuilt ones.
pub fn reconstruct_stream_from<R: Read + Send, W: Write + Send>(
    codec: *const rsgpu_codec,
    readers: &mut [Option<R>],
    writers: &mut [Option<W>],
    opts: &StreamOptions,
) -> Result<(), StreamError> {
    let has_reader: Vec<u8> = readers.iter().map(|r| r.is_some() as u8).collect();
    let has_writer: Vec<u8> = writers.iter().map(|w| w.is_some() as u8).collect();
    let s: Streams<R, W> = Streams::new(
        readers.iter_mut().map(|r| r.as_mut().map_or(std::ptr::null_mut(), |r| r as *mut R)).collect(),
        writers.iter_mut().map(|w| w.as_mut().map_or(std::ptr::null_mut(), |w| w as *mut W)).collect(),
    );
    let o = options(opts);
    let mut err = rsgpu_stream_error::default();
    let rc = unsafe {
        rsgpu_reconstruct_stream_from(codec, readers.len(), writers.len(), has_reader.as_ptr(),
                                      has_writer.as_ptr(), read_cb::<R, W>, write_cb::<R, W>, s.user(), &o,
                                      &mut err)
    };
    result(rc, &err, &s, 0)
}

/// What `grow_cb` reaches through `user`: the caller's cursors.  The library asks
/// for room for one shard at a time per thread (different shards concurrently in
/// parallel I/O), so each cursor is touched by one thread at a time.
struct Grow {
    shards: *mut io::Cursor<Vec<u8>>,
}

/// Makes `vec->cap >= need` for shard `shard` through the cursor's own `Vec`: the
/// bytes the library appended so far become the vector's length, the reservation
/// grows it, and the new buffer goes back to the library.
unsafe extern "C" fn grow_cb(user: *mut c_void, shard: usize, v: *mut rsgpu_vec, need: usize) -> i32 {
    let g = &*(user as *const Grow);
    let r = catch_unwind(AssertUnwindSafe(|| {
        let v = &mut *v;
        let bytes = (*g.shards.add(shard)).get_mut();
        bytes.set_len(v.len);
        bytes.reserve_exact(need - v.len);
        v.data = bytes.as_mut_ptr();
        v.cap = bytes.capacity();
    }));
    if r.is_ok() { 0 } else { -1 }
}

/// `ReedSolomon::reconstruct_stream` (stream.rs:839-900): non-empty cursors are
/// present and are read from position 0; empty ones receive the rebuilt bytes,
/// appended through their `Vec` with the position unmoved.  On error the bytes
/// already appended stay, as the reference's writers keep earlier blocks.
pub fn reconstruct_stream(
    codec: *const rsgpu_codec,
    shards: &mut [io::Cursor<Vec<u8>>],
    opts: &StreamOptions,
) -> Result<(), StreamError> {
    let mut vecs: Vec<rsgpu_vec> = shards
        .iter_mut()
        .map(|c| {
            let pos = c.position() as usize;
            let bytes = c.get_mut();
            rsgpu_vec { data: bytes.as_mut_ptr(), len: bytes.len(), cap: bytes.capacity(), pos }
        })
        .collect();
    let present: Vec<bool> = vecs.iter().map(|v| v.len != 0).collect();
    let g = Grow { shards: shards.as_mut_ptr() };
    let o = options(opts);
    let mut err = rsgpu_stream_error::default();
    let rc = unsafe {
        rsgpu_reconstruct_stream(codec, vecs.as_mut_ptr(), vecs.len(), Some(grow_cb),
                                 &g as *const Grow as *mut c_void, &o, &mut err)
    };
    for ((c, v), p) in shards.iter_mut().zip(vecs.iter()).zip(present) {
        if p {
            c.set_position(v.pos as u64);
        } else {
            // SAFETY: the library wrote v.len initialised bytes into this vector's
            // buffer (grow_cb keeps data and cap in step with the Vec).
            unsafe { c.get_mut().set_len(v.len) };
        }
    }
    let none: Streams<io::Empty, io::Sink> = Streams::new(Vec::new(), Vec::new());
    result(rc, &err, &none, 0)
}
