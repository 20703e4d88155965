//! Bindings of `libscatrs_hip` (declared in `include/scatrs_hip.h`): the MI355X library of
//! scatrs' size-distribution ensemble path over homogeneous and layered spheres.
//!
//! * [`ffi`]: the C ABI as `extern "C"` declarations, one per exported symbol.
//! * [`Context`]: a safe owner of a library context (one per host thread, the reference's
//!   one-workspace-per-thread contract, `scatrs/src/lib.rs:186-190`), whose calls return
//!   [`Result`] with the library's status mapped onto the reference's error enum
//!   (`scatrs/src/error.rs:14-32`).
//!
//! Every `d_*` pointer is device memory on the context's GPU (allocated by the caller, for
//! example with `hipMalloc`); host arguments are marked as such. This crate is not built in the
//! container that produced the library (it has no Rust toolchain): it is the binding a scatrs
//! maintainer adds, compiled where cargo and ROCm are both installed.

#![allow(clippy::too_many_arguments)]

use core::ffi::{c_char, c_int, c_void};

/// Status codes of every entry point.
pub const SCATRS_OK: c_int = 0;
pub const SCATRS_INVALID_INPUT: c_int = 1;
pub const SCATRS_OUTSIDE_VALIDATED_DOMAIN: c_int = 2;
pub const SCATRS_NOT_CONVERGED: c_int = 3;
pub const SCATRS_DEVICE_ERROR: c_int = 4;
/// The ABI version this crate was written against (`scatrs_abi_version()`).
pub const SCATRS_ABI_VERSION: c_int = 3;

/// Launch-parameter keys of scatrs_set_tuning.
pub const SCATRS_TUNE_SPHERE_WAVES: c_int = 0;
pub const SCATRS_TUNE_NODE_WAVES: c_int = 1;
pub const SCATRS_TUNE_CHUNK_SHARE: c_int = 2;
pub const SCATRS_TUNE_SPLIT_BLOCKS: c_int = 3;

/// Distribution kinds of the ensemble entry points (`scatrs/src/ensemble/distribution.rs:15-68`).
pub const SCATRS_LOGNORMAL: i32 = 0;
pub const SCATRS_GAMMA: i32 = 1;
pub const SCATRS_MODIFIED_GAMMA: i32 = 2;
pub const SCATRS_POWER_LAW: i32 = 3;
pub const SCATRS_MONODISPERSE: i32 = 4;
pub const SCATRS_SEGMENT: i32 = 5;

pub mod ffi {
    use super::*;

    /// An opaque library context.
    #[repr(C)]
    pub struct ScatrsCtx {
        _p: [u8; 0],
    }

    /// An opaque captured ensemble call (scatrs_ensemble_fixed_graph).
    #[repr(C)]
    pub struct ScatrsGraph {
        _p: [u8; 0],
    }

    extern "C" {
        pub fn scatrs_abi_version() -> c_int;
        pub fn scatrs_create(device: c_int, stream: *mut c_void, out: *mut *mut ScatrsCtx)
            -> c_int;
        pub fn scatrs_destroy(ctx: *mut ScatrsCtx) -> c_int;
        pub fn scatrs_set_stream(ctx: *mut ScatrsCtx, stream: *mut c_void) -> c_int;
        pub fn scatrs_last_error(ctx: *const ScatrsCtx) -> *const c_char;
        pub fn scatrs_set_profiling(ctx: *mut ScatrsCtx, on: c_int) -> c_int;
        /// ms_out: host array of `cap` doubles.
        pub fn scatrs_kernel_times(ctx: *mut ScatrsCtx, ms_out: *mut f64, cap: c_int) -> c_int;
        pub fn scatrs_set_angular_kernel(ctx: *mut ScatrsCtx, kind: c_int) -> c_int;
        pub fn scatrs_set_tuning(ctx: *mut ScatrsCtx, key: c_int, value: i64) -> c_int;
        pub fn scatrs_get_tuning(ctx: *const ScatrsCtx, key: c_int) -> i64;
        /// mu: host array of n_mu cosines.
        pub fn scatrs_set_phase_cosines(ctx: *mut ScatrsCtx, mu: *const f64, n_mu: i64) -> c_int;
        pub fn scatrs_default_n_max(x: f64) -> i64;

        pub fn scatrs_mie_efficiencies(ctx: *mut ScatrsCtx, d_x: *const f64, d_m: *const f64,
            len: i64, n_max: i64, allow_unvalidated: c_int, d_values: *mut f64,
            d_jacobian: *mut f64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_mie_efficiencies_noalloc(ctx: *mut ScatrsCtx, d_x: *const f64,
            d_m: *const f64, len: i64, n_max: i64, allow_unvalidated: c_int,
            d_values: *mut f64, d_jacobian: *mut f64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_mie_efficiencies_host(ctx: *mut ScatrsCtx, x: *const f64, m: *const f64,
            len: i64, n_max: i64, allow_unvalidated: c_int, values: *mut f64,
            jacobian: *mut f64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_mie_terms(ctx: *mut ScatrsCtx, d_x: *const f64, d_m: *const f64, len: i64,
            n_max: i64, allow_unvalidated: c_int, max_terms: *mut i64, total_terms: *mut i64)
            -> c_int;
        pub fn scatrs_mie_coefficients(ctx: *mut ScatrsCtx, d_x: *const f64, d_m: *const f64,
            len: i64, n_max: i64, allow_unvalidated: c_int, width: i64, d_terms: *mut i64,
            d_a: *mut f64, d_b: *mut f64, d_derivatives: *mut f64, total_terms: *mut i64)
            -> c_int;
        pub fn scatrs_layered_terms(ctx: *mut ScatrsCtx, d_x: *const f64, d_m: *const f64,
            count: i64, layers: c_int, n_max: i64, allow_unvalidated: c_int, derivatives: c_int,
            max_terms: *mut i64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_layered_efficiencies(ctx: *mut ScatrsCtx, d_x: *const f64,
            d_m: *const f64, count: i64, layers: c_int, n_max: i64, allow_unvalidated: c_int,
            d_values: *mut f64, d_jacobian: *mut f64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_layered_coefficients(ctx: *mut ScatrsCtx, d_x: *const f64,
            d_m: *const f64, count: i64, layers: c_int, n_max: i64, allow_unvalidated: c_int,
            width: i64, d_terms: *mut i64, d_a: *mut f64, d_b: *mut f64,
            d_derivatives: *mut f64, total_terms: *mut i64) -> c_int;

        pub fn scatrs_ensemble_fixed(ctx: *mut ScatrsCtx, count: i64, d_kind: *const i32,
            d_p0: *const f64, d_p1: *const f64, d_p2: *const f64, d_wavelength: *const f64,
            layers: c_int, d_fractions: *const f64, d_m: *const f64, n_radius: i64,
            n_angles: i64, order: i64, normalized: c_int, allow_unvalidated: c_int,
            d_scalars: *mut f64, d_greek: *mut f64, d_phase: *mut f64, total_terms: *mut i64)
            -> c_int;
        pub fn scatrs_ensemble_fixed_vjp(ctx: *mut ScatrsCtx, count: i64, d_kind: *const i32,
            d_p0: *const f64, d_p1: *const f64, d_p2: *const f64, d_wavelength: *const f64,
            layers: c_int, d_fractions: *const f64, d_m: *const f64, n_radius: i64,
            n_angles: i64, order: i64, normalized: c_int, fixed_range: c_int,
            allow_unvalidated: c_int, d_bar_scalars: *const f64, d_bar_greek: *const f64,
            d_bar_phase: *const f64, d_grad: *mut f64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_ensemble_fixed_graph(ctx: *mut ScatrsCtx, count: i64, d_kind: *const i32,
            d_p0: *const f64, d_p1: *const f64, d_p2: *const f64, d_wavelength: *const f64,
            layers: c_int, d_fractions: *const f64, d_m: *const f64, n_radius: i64,
            n_angles: i64, order: i64, normalized: c_int, allow_unvalidated: c_int,
            d_scalars: *mut f64, d_greek: *mut f64, d_phase: *mut f64, total_terms: *mut i64,
            out: *mut *mut ScatrsGraph) -> c_int;
        pub fn scatrs_graph_launch(g: *mut ScatrsGraph, stream: *mut c_void) -> c_int;
        pub fn scatrs_graph_status(g: *mut ScatrsGraph, stream: *mut c_void) -> c_int;
        pub fn scatrs_graph_last_error(g: *const ScatrsGraph) -> *const c_char;
        pub fn scatrs_graph_destroy(g: *mut ScatrsGraph) -> c_int;
        pub fn scatrs_ensemble_nodes(ctx: *mut ScatrsCtx, count: i64, n_nodes: i64,
            d_nodes: *const f64, d_wavelength: *const f64, layers: c_int,
            d_fractions: *const f64, d_m: *const f64, n_angles: i64, order: i64,
            normalized: c_int, allow_unvalidated: c_int, d_scalars: *mut f64,
            d_greek: *mut f64, d_phase: *mut f64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_ensemble_nodes_tangent(ctx: *mut ScatrsCtx, count: i64, n_nodes: i64,
            d_nodes: *const f64, d_wavelength: *const f64, layers: c_int,
            d_fractions: *const f64, d_m: *const f64, n_angles: i64, order: i64,
            direction: c_int, sign: c_int, allow_unvalidated: c_int, d_scalars: *mut f64,
            d_greek: *mut f64, d_phase: *mut f64, total_terms: *mut i64) -> c_int;
        pub fn scatrs_ensemble_nodes_vjp(ctx: *mut ScatrsCtx, count: i64, n_nodes: i64,
            d_nodes: *const f64, d_dlnw: *const f64, d_wavelength: *const f64, layers: c_int,
            d_fractions: *const f64, d_m: *const f64, n_angles: i64, order: i64,
            normalized: c_int, allow_unvalidated: c_int, d_bar_scalars: *const f64,
            d_bar_greek: *const f64, d_bar_phase: *const f64, d_grad: *mut f64,
            total_terms: *mut i64) -> c_int;

        /// Host-only helpers: p is a host array of 3 parameters.
        pub fn scatrs_size_range(kind: i32, p: *const f64, t_lo: *mut f64, t_hi: *mut f64)
            -> c_int;
        pub fn scatrs_size_message(kind: i32, p: *const f64, buf: *mut c_char, cap: i64)
            -> c_int;
        pub fn scatrs_size_log_density(n: i64, kind: *const i32, p: *const f64, r: *const f64,
            ln_u: *mut f64, d_ln_u: *mut f64) -> c_int;
        /// The library's n-point Gauss-Legendre rule (1 <= n <= 65536) into host arrays.
        pub fn scatrs_gauss_legendre(n: i64, nodes: *mut f64, weights: *mut f64) -> c_int;
    }
}

/// The library's failures, one variant per status (the reference's `scatrs::Error` variants,
/// `scatrs/src/error.rs:14-32`, plus the device's own failures).
#[derive(Debug, Clone, PartialEq)]
pub enum Error {
    /// `SCATRS_INVALID_INPUT`: nothing was computed.
    InvalidInput(String),
    /// `SCATRS_OUTSIDE_VALIDATED_DOMAIN`: valid inputs outside the validated domain.
    OutsideValidatedDomain(String),
    /// `SCATRS_NOT_CONVERGED`: an iteration missed the accuracy it promises.
    NotConverged(String),
    /// `SCATRS_DEVICE_ERROR`: a HIP call failed (allocation, launch, copy).
    Device(String),
}

impl core::fmt::Display for Error {
    fn fmt(&self, f: &mut core::fmt::Formatter<'_>) -> core::fmt::Result {
        match self {
            Error::InvalidInput(m) => write!(f, "invalid input: {m}"),
            Error::OutsideValidatedDomain(m) => write!(f, "outside validated domain: {m}"),
            Error::NotConverged(m) => write!(f, "not converged: {m}"),
            Error::Device(m) => write!(f, "device error: {m}"),
        }
    }
}

impl std::error::Error for Error {}

/// With the `scatrs-error` feature the library's errors convert to the reference's own enum.
/// Its `NotConverged` carries numbers, not a message: the library reports none there (no
/// entry point of this path fails to converge: a size integral that misses its tolerance is a
/// result with `converged = false`, `scatrs/src/error.rs:11-12`), so they are NaN.
#[cfg(feature = "scatrs-error")]
impl From<Error> for scatrs::Error {
    fn from(e: Error) -> Self {
        match e {
            Error::InvalidInput(m) => scatrs::Error::InvalidInput(m),
            Error::OutsideValidatedDomain(m) => scatrs::Error::OutsideValidatedDomain(m),
            Error::NotConverged(_) => {
                scatrs::Error::NotConverged { achieved: f64::NAN, requested: f64::NAN }
            }
            Error::Device(m) => scatrs::Error::InvalidInput(format!("device: {m}")),
        }
    }
}

pub type Result<T> = core::result::Result<T, Error>;

/// An owned library context on one GPU, issuing work on one HIP stream.
pub struct Context {
    raw: *mut ffi::ScatrsCtx,
}

// A context may move between threads but not be shared: the library's contexts are not
// thread-safe.
unsafe impl Send for Context {}

impl Context {
    /// A context on `device` issuing work on `stream` (a `hipStream_t`; null for the null
    /// stream). Fails if the loaded library's ABI differs from this crate's.
    pub fn new(device: i32, stream: *mut c_void) -> Result<Self> {
        let abi = unsafe { ffi::scatrs_abi_version() };
        if abi != SCATRS_ABI_VERSION {
            return Err(Error::Device(format!(
                "libscatrs_hip has ABI version {abi}, this binding expects {SCATRS_ABI_VERSION}"
            )));
        }
        let mut raw = core::ptr::null_mut();
        let rc = unsafe { ffi::scatrs_create(device, stream, &mut raw) };
        if rc != SCATRS_OK || raw.is_null() {
            return Err(Error::Device(format!("scatrs_create failed on device {device} ({rc})")));
        }
        Ok(Self { raw })
    }

    pub fn as_ptr(&mut self) -> *mut ffi::ScatrsCtx {
        self.raw
    }

    fn message(&self) -> String {
        unsafe {
            let p = ffi::scatrs_last_error(self.raw);
            if p.is_null() {
                String::new()
            } else {
                std::ffi::CStr::from_ptr(p).to_string_lossy().into_owned()
            }
        }
    }

    /// Maps a status onto [`Error`] with the context's message.
    pub fn check(&self, rc: c_int) -> Result<()> {
        match rc {
            SCATRS_OK => Ok(()),
            SCATRS_INVALID_INPUT => Err(Error::InvalidInput(self.message())),
            SCATRS_OUTSIDE_VALIDATED_DOMAIN => Err(Error::OutsideValidatedDomain(self.message())),
            SCATRS_NOT_CONVERGED => Err(Error::NotConverged(self.message())),
            _ => Err(Error::Device(self.message())),
        }
    }

    /// Efficiencies (5, len) and optionally their Jacobian (3, 5, len) of `len` spheres; returns
    /// the number of (particle, order) terms. The reference's `_scatrs.mie_efficiencies`
    /// batch (`scatrs-py/src/mie.rs:303-364`).
    ///
    /// # Safety
    /// The pointers are device memory of the stated sizes on this context's GPU.
    pub unsafe fn mie_efficiencies(&mut self, d_x: *const f64, d_m: *const f64, len: i64,
                                   n_max: Option<i64>, allow_unvalidated: bool,
                                   d_values: *mut f64, d_jacobian: *mut f64) -> Result<i64> {
        let mut terms = 0i64;
        let rc = ffi::scatrs_mie_efficiencies(self.raw, d_x, d_m, len, n_max.unwrap_or(0),
                                              allow_unvalidated as c_int, d_values, d_jacobian,
                                              &mut terms);
        self.check(rc).map(|_| terms)
    }

    /// [`Context::mie_efficiencies`] on host slices, as `_scatrs.mie_efficiencies` takes numpy
    /// arrays: x (len), m (len complex as re, im pairs), values (5 len), jacobian (15 len) if
    /// given. One synchronization; the checks and the ranking run on the host.
    pub fn mie_efficiencies_host(&mut self, x: &[f64], m: &[f64], n_max: Option<i64>,
                                 allow_unvalidated: bool, values: &mut [f64],
                                 jacobian: Option<&mut [f64]>) -> Result<i64> {
        let len = x.len();
        assert!(m.len() == 2 * len && values.len() == 5 * len);
        let jac = match jacobian {
            Some(j) => {
                assert!(j.len() == 15 * len);
                j.as_mut_ptr()
            }
            None => std::ptr::null_mut(),
        };
        let mut terms = 0i64;
        // Safety: the slices hold the sizes asserted above; the library reads and writes only
        // those elements, before it returns.
        let rc = unsafe {
            ffi::scatrs_mie_efficiencies_host(self.raw, x.as_ptr(), m.as_ptr(), len as i64,
                                              n_max.unwrap_or(0), allow_unvalidated as c_int,
                                              values.as_mut_ptr(), jac, &mut terms)
        };
        self.check(rc).map(|_| terms)
    }

    /// Layered efficiencies (5, count) and optionally the per-layer Jacobian
    /// (layers, 3, 5, count) (`scatrs-py/src/layered.rs:123-193`).
    ///
    /// # Safety
    /// As [`Context::mie_efficiencies`]; d_x is (count, layers) and d_m (count, layers)
    /// complex128.
    pub unsafe fn layered_efficiencies(&mut self, d_x: *const f64, d_m: *const f64, count: i64,
                                       layers: i32, n_max: Option<i64>,
                                       allow_unvalidated: bool, d_values: *mut f64,
                                       d_jacobian: *mut f64) -> Result<i64> {
        let mut terms = 0i64;
        let rc = ffi::scatrs_layered_efficiencies(self.raw, d_x, d_m, count, layers,
                                                  n_max.unwrap_or(0), allow_unvalidated as c_int,
                                                  d_values, d_jacobian, &mut terms);
        self.check(rc).map(|_| terms)
    }

    /// The largest and total N of `len` spheres, for sizing coefficient outputs.
    ///
    /// # Safety
    /// As [`Context::mie_efficiencies`].
    pub unsafe fn mie_terms(&mut self, d_x: *const f64, d_m: *const f64, len: i64,
                            n_max: Option<i64>, allow_unvalidated: bool) -> Result<(i64, i64)> {
        let (mut most, mut total) = (0i64, 0i64);
        let rc = ffi::scatrs_mie_terms(self.raw, d_x, d_m, len, n_max.unwrap_or(0),
                                       allow_unvalidated as c_int, &mut most, &mut total);
        self.check(rc).map(|_| (most, total))
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe {
            ffi::scatrs_destroy(self.raw);
        }
    }
}
