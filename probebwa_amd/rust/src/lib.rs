//! `extern "C"` bindings of libprobebwa_amd (include/probebwa_amd.h): the MI355X
//! (gfx950) implementation of probebwa's `map` path.  Every function returns 0 on
//! success and a negative status (`PBW_ERR_*`) on error, with the message from
//! `pbw_last_error` (per context) or `pbw_global_error`.
//!
//! The types mirror the C structs field for field (`#[repr(C)]`); the opaque
//! handles are only ever used behind pointers.  `Genome` and `Context` below are
//! thin owners that free their handle on drop.
#![allow(non_camel_case_types)]

use std::ffi::{CStr, CString};
use std::os::raw::{c_char, c_int};

pub const PBW_OK: c_int = 0;
pub const PBW_ERR_ARG: c_int = -1;
pub const PBW_ERR_HIP: c_int = -2;
pub const PBW_ERR_LIMIT: c_int = -3;
pub const PBW_ERR_IO: c_int = -4;
pub const PBW_ERR_MISMATCH: c_int = -5;
pub const PBW_TUNE_MAX_TRIALS: usize = 48;
pub const PBW_MLP_INPUTS: usize = 32;
pub const PBW_MLP_PARAMS: usize = 20865;
pub const PBW_LABEL_LOGLK: c_int = 0;
pub const PBW_LABEL_SCORE: c_int = 1;

#[repr(C)]
pub struct pbw_genome {
    _p: [u8; 0],
}
#[repr(C)]
pub struct pbw_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct pbw_device_reads {
    _p: [u8; 0],
}
#[repr(C)]
pub struct pbw_fastx {
    _p: [u8; 0],
}

/// One output record (the AlignmentRecord fields SAM carries, src/types.rs:180-203).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_record {
    pub pos: u64,
    pub mate_pos: u64,
    pub tlen: i64,
    pub log_lk: f64,
    pub posterior: f64,
    pub contig: i32,
    pub mate_contig: i32,
    pub cigar_off: u32,
    pub cigar_n: u32,
    pub flag: u16,
    pub mapq: u8,
    pub strand: u8,
    pub seq_rc: u8,
    pub pad: [u8; 3],
}

/// MapperOptions (src/types.rs:207-249) plus the device batch size.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_options {
    pub substitution_rate: f64,
    pub insert_size_mean: f64,
    pub insert_size_sd: f64,
    pub sv_prior_phred: f64,
    pub max_indel_len: u32,
    pub gap_open_phred: u32,
    pub gap_extend_phred: u32,
    pub batch_reads: u32,
}

/// Reads: bases upper-cased, qualities offset-subtracted.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct pbw_read_batch {
    pub n: u64,
    pub seq: *const u8,
    pub seq_off: *const u64,
    pub qual: *const u8,
    pub qual_off: *const u64,
}

#[repr(C)]
#[derive(Debug)]
pub struct pbw_record_batch {
    pub n: u64,
    pub rec: *mut pbw_record,
    pub cigar: *mut u32,
    pub n_cigar: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_insert_model {
    pub mean: f64,
    pub sd: f64,
    pub n_observed: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct pbw_align_job {
    pub read: *const u8,
    pub qual: *const u8,
    pub reference: *const u8,
    pub m: u32,
    pub qlen: u32,
    pub n: u32,
    pub band: u32,
    pub hint: i64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_align_result {
    pub score: i32,
    pub ref_offset: u32,
    pub cigar_off: u32,
    pub cigar_n: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_stats {
    pub reads: u64,
    pub lookups: u64,
    pub positions: u64,
    pub filtered_hits: u64,
    pub dp_jobs: u64,
    pub dp_cells: u64,
    pub wide_jobs: u64,
    pub wide_cells: u64,
    pub seed_overflow_reads: u64,
    pub tier2_cluster_reads: u64,
    pub generic_dp_jobs: u64,
    pub band_dp_jobs: u64,
    pub ms_seed: f64,
    pub ms_cluster: f64,
    pub ms_dp: f64,
    pub ms_mapq: f64,
    pub ms_pair: f64,
    pub ms_total: f64,
    pub ms_k_seed: f64,
    pub ms_k_dp: f64,
    pub n_k_seed: u64,
    pub n_k_dp: u64,
    pub pf_jobs: u64,
    pub pf_jobs_aligned: u64,
    pub pf_fallback_reads: u64,
    pub pf_audit_reads: u64,
    pub pf_audit_mismatches: u64,
    pub ms_pf: f64,
    pub ms_k_cluster: f64,
    pub n_k_cluster: u64,
    pub mapq_near_boundary: u64,
    pub ms_k_wide: f64,
    pub n_k_wide: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_tune {
    pub batch_reads: u32,
    pub dp_grid: u32,
    pub split_len: u32,
    pub long_waves: u32,
    pub dp_pack16: u32,
    pub seed_pool: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct pbw_autotune_report {
    pub best: pbw_tune,
    pub best_ms: f64,
    pub trials: u32,
    pub tried: [pbw_tune; PBW_TUNE_MAX_TRIALS],
    pub ms: [f64; PBW_TUNE_MAX_TRIALS],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_train_options {
    pub epochs: u32,
    pub batch: u32,
    pub lr: f32,
    pub beta1: f32,
    pub beta2: f32,
    pub eps: f32,
    pub pos_weight: f32,
    pub seed: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_prefilter_options {
    pub enabled: i32,
    pub threshold: f32,
    pub top_k: u32,
    pub mapq_fallback: u32,
    pub close_nats: f32,
    pub audit_period: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_prefilter_report {
    pub rows: u64,
    pub positives: u64,
    pub val_rows: u64,
    pub val_positives: u64,
    pub threshold: f32,
    pub val_recall: f64,
    pub val_skip: f64,
    pub loss_first: f64,
    pub loss_last: f64,
}

/// The map subcommand's read and output flags (src/main.rs:40-111).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct pbw_io_options {
    pub input_format: c_int,
    pub quality_offset: u32,
    pub adapter: *const c_char,
    pub label_prefix: *const c_char,
    pub command_line: *const c_char,
    pub read_group: *const c_char,
    pub output_format: c_int,
    pub index: c_int,
    pub raw_qual: c_int,
    pub threads: u32,
    pub batch_reads: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct pbw_io_report {
    pub reads: u64,
    pub records: u64,
    pub bytes_in: u64,
    pub bytes_out: u64,
    pub ms_parse: f64,
    pub ms_map: f64,
    pub ms_write: f64,
    pub ms_total: f64,
}

extern "C" {
    // genome: GenomeIndexBuilder / GenomeIndex (src/index/genome.rs)
    pub fn pbw_genome_create(out: *mut *mut pbw_genome) -> c_int;
    pub fn pbw_genome_add_contig(g: *mut pbw_genome, name: *const c_char, seq: *const u8, len: u64) -> c_int;
    pub fn pbw_genome_add_fasta(g: *mut pbw_genome, path: *const c_char) -> c_int;
    pub fn pbw_genome_n_contigs(g: *const pbw_genome) -> c_int;
    pub fn pbw_genome_contig_name(g: *const pbw_genome, cid: c_int) -> *const c_char;
    pub fn pbw_genome_contig_length(g: *const pbw_genome, cid: c_int) -> u64;
    pub fn pbw_genome_destroy(g: *mut pbw_genome);
    pub fn pbw_genome_get_sequence(g: *const pbw_genome, cid: c_int, start: u64, end: u64, out: *mut u8) -> u64;
    pub fn pbw_genome_load_stidx(path: *const c_char, out: *mut *mut pbw_genome) -> c_int;
    pub fn pbw_genome_save_stidx(g: *const pbw_genome, species: *const c_char, assembly: *const c_char,
                                 path: *const c_char) -> c_int;

    // context: index build (or HashTable::load) and upload, SingleEndMapper::new
    pub fn pbw_default_options(o: *mut pbw_options);
    pub fn pbw_ctx_create(g: *const pbw_genome, o: *const pbw_options, device: c_int, out: *mut *mut pbw_ctx)
        -> c_int;
    pub fn pbw_ctx_create_from_sthash(g: *const pbw_genome, sthash_path: *const c_char, o: *const pbw_options,
                                      device: c_int, out: *mut *mut pbw_ctx) -> c_int;
    pub fn pbw_index_save_sthash(ctx: *mut pbw_ctx, path: *const c_char) -> c_int;
    pub fn pbw_ctx_genome(ctx: *const pbw_ctx) -> *const pbw_genome;
    pub fn pbw_ctx_get_options(ctx: *const pbw_ctx, out: *mut pbw_options) -> c_int;
    pub fn pbw_last_error(ctx: *const pbw_ctx) -> *const c_char;
    pub fn pbw_global_error() -> *const c_char;
    pub fn pbw_ctx_destroy(ctx: *mut pbw_ctx);

    // mapping
    pub fn pbw_map_se(ctx: *mut pbw_ctx, reads: *const pbw_read_batch, out: *mut pbw_record_batch) -> c_int;
    pub fn pbw_reads_upload(ctx: *mut pbw_ctx, reads: *const pbw_read_batch, out: *mut *mut pbw_device_reads)
        -> c_int;
    pub fn pbw_map_se_device(ctx: *mut pbw_ctx, reads: *const pbw_device_reads, begin: u64, end: u64,
                             out: *mut pbw_record_batch) -> c_int;
    pub fn pbw_reads_free(reads: *mut pbw_device_reads);
    pub fn pbw_tune_get(ctx: *const pbw_ctx, out: *mut pbw_tune) -> c_int;
    pub fn pbw_tune_set(ctx: *mut pbw_ctx, t: *const pbw_tune) -> c_int;
    pub fn pbw_autotune(ctx: *mut pbw_ctx, reads: *const pbw_device_reads, begin: u64, end: u64,
                        out: *mut pbw_autotune_report) -> c_int;
    pub fn pbw_map_pe(ctx: *mut pbw_ctx, m1: *const pbw_read_batch, m2: *const pbw_read_batch,
                      inout: *mut pbw_insert_model, out: *mut pbw_record_batch) -> c_int;
    pub fn pbw_align_batch(ctx: *mut pbw_ctx, gap_open_phred: u32, gap_extend_phred: u32,
                           jobs: *const pbw_align_job, n: u64, results: *mut pbw_align_result, cigar: *mut u32,
                           cigar_cap: u64) -> c_int;
    pub fn pbw_align_batch_noalloc(ctx: *mut pbw_ctx, gap_open_phred: u32, gap_extend_phred: u32,
                                   jobs: *const pbw_align_job, n: u64, results: *mut pbw_align_result,
                                   cigar: *mut u32, cigar_cap: u64) -> c_int;
    pub fn pbw_index_lookup(ctx: *mut pbw_ctx, code: u32, out: *mut u64, cap: u64) -> i64;
    pub fn pbw_index_total_entries(ctx: *const pbw_ctx) -> u64;

    // candidate prefilter network
    pub fn pbw_default_train_options(o: *mut pbw_train_options);
    pub fn pbw_default_prefilter_options(o: *mut pbw_prefilter_options);
    pub fn pbw_prefilter_set(ctx: *mut pbw_ctx, o: *const pbw_prefilter_options) -> c_int;
    pub fn pbw_prefilter_get(ctx: *const pbw_ctx, o: *mut pbw_prefilter_options) -> c_int;
    pub fn pbw_prefilter_dataset(ctx: *mut pbw_ctx, reads: *const pbw_read_batch, label_kind: c_int, margin: f32,
                                 x: *mut u16, y: *mut f32, rank: *mut u8, cap: u64, n: *mut u64) -> c_int;
    pub fn pbw_prefilter_train(ctx: *mut pbw_ctx, reads: *const pbw_read_batch, o: *const pbw_train_options,
                               label_kind: c_int, margin: f32, recall: f32, out: *mut pbw_prefilter_report)
        -> c_int;
    pub fn pbw_mlp_init(ctx: *mut pbw_ctx, seed: u64) -> c_int;
    pub fn pbw_mlp_set_params(ctx: *mut pbw_ctx, theta: *const f32, n: u64) -> c_int;
    pub fn pbw_mlp_get_params(ctx: *mut pbw_ctx, theta: *mut f32, n: u64) -> c_int;
    pub fn pbw_mlp_forward(ctx: *mut pbw_ctx, x_bf16: *const u16, n: u64, logits: *mut f32) -> c_int;
    pub fn pbw_mlp_gradient(ctx: *mut pbw_ctx, x_bf16: *const u16, y: *const f32, n: u64, pos_weight: f32,
                            grad: *mut f32, loss: *mut f64) -> c_int;
    pub fn pbw_mlp_train(ctx: *mut pbw_ctx, x_bf16: *const u16, y: *const f32, n: u64,
                         o: *const pbw_train_options, epoch_loss: *mut f64) -> c_int;

    pub fn pbw_records_release(b: *mut pbw_record_batch);
    pub fn pbw_get_stats(ctx: *const pbw_ctx, out: *mut pbw_stats) -> c_int;

    // output and files
    pub fn pbw_format_sam(g: *const pbw_genome, reads: *const pbw_read_batch, ids: *const c_char,
                          recs: *const pbw_record_batch, command_line: *const c_char, read_group: *const c_char,
                          len: *mut u64) -> *mut c_char;
    pub fn pbw_free_text(t: *mut c_char);
    pub fn pbw_write_bam(g: *const pbw_genome, reads: *const pbw_read_batch, ids: *const c_char,
                         recs: *const pbw_record_batch, command_line: *const c_char, read_group: *const c_char,
                         path: *const c_char, index: c_int, raw_qual: c_int, threads: u32) -> c_int;
    pub fn pbw_fastx_open(path: *const c_char, format: c_int, quality_offset: u32, adapter: *const c_char,
                          out: *mut *mut pbw_fastx) -> c_int;
    pub fn pbw_fastx_next(f: *mut pbw_fastx, max_reads: u64, reads: *mut pbw_read_batch, ids: *mut *const c_char)
        -> c_int;
    pub fn pbw_fastx_bytes_read(f: *const pbw_fastx) -> u64;
    pub fn pbw_fastx_close(f: *mut pbw_fastx);
    pub fn pbw_default_io_options(o: *mut pbw_io_options);
    pub fn pbw_map_files(ctx: *mut pbw_ctx, reads1: *const c_char, reads2: *const c_char, out_path: *const c_char,
                         o: *const pbw_io_options, report: *mut pbw_io_report) -> c_int;
}

/// The library's message for a failed call (`ctx` may be null).
pub fn error_message(ctx: *const pbw_ctx) -> String {
    unsafe {
        let p = if ctx.is_null() { pbw_global_error() } else { pbw_last_error(ctx) };
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}

fn check(rc: c_int, ctx: *const pbw_ctx) -> Result<(), String> {
    if rc == PBW_OK {
        Ok(())
    } else {
        Err(format!("probebwa_amd error {}: {}", rc, error_message(ctx)))
    }
}

/// Owner of a `pbw_genome`.
pub struct Genome(pub *mut pbw_genome);

impl Genome {
    pub fn new() -> Result<Self, String> {
        let mut g = std::ptr::null_mut();
        check(unsafe { pbw_genome_create(&mut g) }, std::ptr::null())?;
        Ok(Genome(g))
    }
    /// GenomeIndex::load (src/index/genome.rs:80-86).
    pub fn load_stidx(path: &str) -> Result<Self, String> {
        let p = CString::new(path).map_err(|e| e.to_string())?;
        let mut g = std::ptr::null_mut();
        check(unsafe { pbw_genome_load_stidx(p.as_ptr(), &mut g) }, std::ptr::null())?;
        Ok(Genome(g))
    }
    pub fn add_contig(&mut self, name: &str, seq: &[u8]) -> Result<(), String> {
        let n = CString::new(name).map_err(|e| e.to_string())?;
        check(unsafe { pbw_genome_add_contig(self.0, n.as_ptr(), seq.as_ptr(), seq.len() as u64) },
              std::ptr::null())
    }
}

impl Drop for Genome {
    fn drop(&mut self) {
        unsafe { pbw_genome_destroy(self.0) }
    }
}

/// Owner of a `pbw_ctx` (device index and stream); the genome must outlive it.
pub struct Context(pub *mut pbw_ctx);

impl Context {
    pub fn new(g: &Genome, o: &pbw_options, device: c_int) -> Result<Self, String> {
        let mut c = std::ptr::null_mut();
        check(unsafe { pbw_ctx_create(g.0, o, device, &mut c) }, std::ptr::null())?;
        Ok(Context(c))
    }
    /// The bucket table read from a `.sthash` (HashTable::load, src/index/hashtable.rs:66-72).
    pub fn from_sthash(g: &Genome, sthash: &str, o: &pbw_options, device: c_int) -> Result<Self, String> {
        let p = CString::new(sthash).map_err(|e| e.to_string())?;
        let mut c = std::ptr::null_mut();
        check(unsafe { pbw_ctx_create_from_sthash(g.0, p.as_ptr(), o, device, &mut c) }, std::ptr::null())?;
        Ok(Context(c))
    }
    /// Single-end mapping (the rayon loop of map_reads_with_format, src/lib.rs:84-104);
    /// `f` sees the records before the library frees them.
    pub fn map_se<T>(&mut self, reads: &pbw_read_batch,
                     f: impl FnOnce(&[pbw_record], &[u32]) -> T) -> Result<T, String> {
        let mut out = pbw_record_batch { n: 0, rec: std::ptr::null_mut(), cigar: std::ptr::null_mut(), n_cigar: 0 };
        check(unsafe { pbw_map_se(self.0, reads, &mut out) }, self.0)?;
        let r = unsafe {
            let rec: &[pbw_record] =
                if out.n == 0 { &[] } else { std::slice::from_raw_parts(out.rec, out.n as usize) };
            let cig: &[u32] =
                if out.n_cigar == 0 { &[] } else { std::slice::from_raw_parts(out.cigar, out.n_cigar as usize) };
            f(rec, cig)
        };
        unsafe { pbw_records_release(&mut out) };
        Ok(r)
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { pbw_ctx_destroy(self.0) }
    }
}
