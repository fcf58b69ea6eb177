;;; fl-gpu.lisp -- femlisp's side of libflgpu.so (include/flgpu.h).
;;;
;;; Load after femlisp (SBCL).  It binds the C-ABI with fl.port:def-function
;;; (src/basic/port.lisp:442-455) in the manner of femlisp's SuperLU binding
;;; (src/alien/superlu.lisp:57-75), defines <gpu-sparse-matrix> and <gpu-vector>
;;; with the matlisp generics the solver path calls, make-iterator methods for
;;; <jacobi> and <sor> on GPU matrices, the flgpu_mg_create binding (the hierarchy
;;; and its parameters are built by the caller), and EXPORT-SYSTEM, which writes a
;;; femlisp system in the layout flgpu_system_read reads.
;;;
;;; No Lisp implementation is installed where this library is built, so this file
;;; has not been executed; the same C-ABI is exercised by the Python host side and
;;; the tests of the repository.

(in-package "FL.ALIEN")

(defparameter *flgpu-library*
  (or (probe-file (fl.start::femlisp-pathname "interface/libflgpu.so"))
      (probe-file "libflgpu.so")))

(defun load-flgpu-library ()
  (load-foreign-library *flgpu-library*))

(when *flgpu-library*
  (load-flgpu-library)
  (pushnew 'load-flgpu-library *foreign-code-loaders*)
  (pushnew :flgpu *features*))

(eval-when (:compile-toplevel :load-toplevel :execute)
  (defconstant +long+ #+sbcl 'sb-alien:long #-sbcl :long))

;;; ---- foreign functions (int status, caller-owned host arrays) ------------

#+flgpu
(progn
  (fl.port:def-function ("flgpu_last_error" flgpu-last-error)
      ((buf (* :char)) (len :int)) :returning :int)
  (fl.port:def-function ("flgpu_ctx_create" flgpu-ctx-create)
      ((device :int) (out (* (* t)))) :returning :int)
  (fl.port:def-function ("flgpu_vec_create" flgpu-vec-create)
      ((ctx (* t)) (n #.+long+) (prec :int) (out (* (* t)))) :returning :int)
  (fl.port:def-function ("flgpu_vec_destroy" flgpu-vec-destroy) ((v (* t))) :returning :int)
  (fl.port:def-function ("flgpu_vec_upload" flgpu-vec-upload)
      ((v (* t)) (host (* :double))) :returning :int)
  (fl.port:def-function ("flgpu_vec_download" flgpu-vec-download)
      ((v (* t)) (host (* :double))) :returning :int)
  (fl.port:def-function ("flgpu_vec_copy" flgpu-vec-copy) ((x (* t)) (y (* t))) :returning :int)
  (fl.port:def-function ("flgpu_vec_axpy" flgpu-vec-axpy)
      ((a :double) (x (* t)) (y (* t))) :returning :int)
  (fl.port:def-function ("flgpu_vec_scal" flgpu-vec-scal) ((a :double) (x (* t))) :returning :int)
  (fl.port:def-function ("flgpu_vec_fill" flgpu-vec-fill) ((x (* t)) (a :double)) :returning :int)
  (fl.port:def-function ("flgpu_vec_dot" flgpu-vec-dot)
      ((x (* t)) (y (* t)) (out (* :double))) :returning :int)
  (fl.port:def-function ("flgpu_csr_upload" flgpu-csr-upload)
      ((ctx (* t)) (nrows :int) (ncols :int) (nnz #.+long+)
       (rowptr (* :int)) (colidx (* :int)) (vals (* :double))
       (prec :int) (out (* (* t))))
    :returning :int)
  (fl.port:def-function ("flgpu_mat_destroy" flgpu-mat-destroy) ((a (* t))) :returning :int)
  (fl.port:def-function ("flgpu_gemm_nn" flgpu-gemm-nn)
      ((alpha :double) (a (* t)) (y (* t)) (beta :double) (x (* t))) :returning :int)
  (fl.port:def-function ("flgpu_gemm_tn" flgpu-gemm-tn)
      ((alpha :double) (a (* t)) (y (* t)) (beta :double) (x (* t))) :returning :int)
  (fl.port:def-function ("flgpu_compute_residual" flgpu-compute-residual)
      ((a (* t)) (x (* t)) (b (* t)) (r (* t))) :returning :int)
  (fl.port:def-function ("flgpu_mat_colour_greedy" flgpu-mat-colour-greedy)
      ((a (* t)) (ncolours (* :int))) :returning :int)
  (fl.port:def-function ("flgpu_jacobi_iterate" flgpu-jacobi-iterate)
      ((a (* t)) (damp :double) (x (* t)) (r (* t))) :returning :int)
  (fl.port:def-function ("flgpu_sor_iterate" flgpu-sor-iterate)
      ((a (* t)) (omega :double) (backwards :int) (x (* t)) (b (* t)) (r (* t)))
    :returning :int)
  (fl.port:def-function ("flgpu_lattice_colouring" flgpu-lattice-colouring)
      ((a (* t)) (dim :int) (order :int) (n :int)) :returning :int)
  (fl.port:def-function ("flgpu_block_iterate" flgpu-block-iterate)
      ((a (* t)) (kind :int) (omega :double) (backwards :int) (x (* t)) (b (* t)) (r (* t)))
    :returning :int)
  (fl.port:def-function ("flgpu_mg_create" flgpu-mg-create)
      ((ctx (* t)) (nlevels :int) (a-vec (* (* t))) (i-vec (* (* t)))
       (params (* t)) (out (* (* t))))
    :returning :int)
  (fl.port:def-function ("flgpu_mg_apply" flgpu-mg-apply)
      ((mg (* t)) (x (* t)) (b (* t)) (r (* t))) :returning :int)
  (fl.port:def-function ("flgpu_system_read" flgpu-system-read)
      ((ctx (* t)) (path sb-alien:c-string) (prec :int) (a (* (* t))) (rhs (* (* t))))
    :returning :int))

(defun flgpu-check (status)
  "femlisp's reaction to a failed foreign solver call (compressed.lisp:327-333)."
  (unless (zerop status)
    (let ((buf (make-array 1024 :element-type '(unsigned-byte 8))))
      (fl.port:foreign-call #'flgpu-last-error buf 1024)
      (error "flgpu status ~D: ~A" status
             (map 'string #'code-char (subseq buf 0 (position 0 buf)))))))

(defun flgpu-out-handle (function &rest args)
  "Calls FUNCTION with ARGS and a one-slot output box; returns the handle."
  (let ((box (make-array 1 :element-type '(unsigned-byte 64))))
    (flgpu-check (apply #'fl.port:foreign-call function (append args (list box))))
    (sb-sys:int-sap (aref box 0))))

(defvar *flgpu-ctx* nil "The context (device 0, one HIP stream).")

(defun flgpu-ctx ()
  (or *flgpu-ctx* (setq *flgpu-ctx* (flgpu-out-handle #'flgpu-ctx-create 0))))

;;; ---- matrices and vectors ------------------------------------------------

(defclass <gpu-vector> ()
  ((handle :initarg :handle :reader gpu-handle)
   (n :initarg :n :reader gpu-size)))

(defun gpu-vector (n &optional (prec 0))
  (make-instance '<gpu-vector> :n n
                 :handle (flgpu-out-handle #'flgpu-vec-create (flgpu-ctx) n prec)))

(defun gpu-vector-from (double-vector)
  (let ((v (gpu-vector (length double-vector))))
    (flgpu-check (fl.port:foreign-call #'flgpu-vec-upload (gpu-handle v) double-vector))
    v))

(defun gpu-vector-to-host (v)
  (let ((out (make-array (gpu-size v) :element-type 'double-float)))
    (flgpu-check (fl.port:foreign-call #'flgpu-vec-download (gpu-handle v) out))
    out))

(defmethod copy! ((x <gpu-vector>) (y <gpu-vector>))
  (flgpu-check (flgpu-vec-copy (gpu-handle x) (gpu-handle y))) y)
(defmethod axpy! (a (x <gpu-vector>) (y <gpu-vector>))
  (flgpu-check (flgpu-vec-axpy (float a 1d0) (gpu-handle x) (gpu-handle y))) y)
(defmethod scal! (a (x <gpu-vector>))
  (flgpu-check (flgpu-vec-scal (float a 1d0) (gpu-handle x))) x)
(defmethod x<-0 ((x <gpu-vector>))
  (flgpu-check (flgpu-vec-fill (gpu-handle x) 0d0)) x)
(defmethod dot ((x <gpu-vector>) (y <gpu-vector>))
  (let ((out (make-array 1 :element-type 'double-float)))
    (flgpu-check (fl.port:foreign-call #'flgpu-vec-dot (gpu-handle x) (gpu-handle y) out))
    (aref out 0)))

(defclass <gpu-sparse-matrix> (<matrix>)
  ((handle :initarg :handle :reader gpu-handle)
   (nrows :initarg :nrows :reader nrows)
   (ncols :initarg :ncols :reader ncols)
   (keys :initarg :keys :reader gpu-keys :documentation "row keys in DOF-number order")
   (lattice :initarg :lattice :initform nil :reader gpu-lattice
            :documentation "(dim order n) when the keys number that unit-cube lattice")))

(defun canonical-keys (A)
  "The row keys of A in sort-lexicographically order of their DOF midpoints
(src/mesh/meshgen.lisp:214-217): on a structured lattice this is the
library's canonical numbering."
  (coerce (fl.mesh::sort-lexicographically (copy-list (row-keys A))) 'vector))

(defun sparse-matrix->gpu (A &key (keys (canonical-keys A)) (prec 0) lattice)
  "Uploads A: the CCS of A^T (sparse-matrix->ccs, src/matlisp/sparselu.lisp:235)
is the CSR of A."
  (let ((ccs (sparse-matrix->ccs (transpose A) :keys keys)))
    (with-slots (sizes starts indices) (pattern ccs)
      (let ((n (aref sizes 0)))
        (make-instance '<gpu-sparse-matrix>
                       :nrows n :ncols (aref sizes 1) :keys keys :lattice lattice
                       :handle (flgpu-out-handle #'flgpu-csr-upload (flgpu-ctx)
                                                 n (aref sizes 1) (aref starts n)
                                                 starts indices (store ccs) prec))))))

(defmethod make-domain-vector-for ((A <gpu-sparse-matrix>) &optional multiplicity)
  (declare (ignore multiplicity))
  (gpu-vector (ncols A)))
(defmethod make-image-vector-for ((A <gpu-sparse-matrix>) &optional multiplicity)
  (declare (ignore multiplicity))
  (gpu-vector (nrows A)))

(defmethod gemm-nn! (alpha (A <gpu-sparse-matrix>) (y <gpu-vector>) beta (x <gpu-vector>))
  (flgpu-check (flgpu-gemm-nn (float alpha 1d0) (gpu-handle A) (gpu-handle y)
                              (float beta 1d0) (gpu-handle x)))
  x)
(defmethod gemm-tn! (alpha (A <gpu-sparse-matrix>) (y <gpu-vector>) beta (x <gpu-vector>))
  (flgpu-check (flgpu-gemm-tn (float alpha 1d0) (gpu-handle A) (gpu-handle y)
                              (float beta 1d0) (gpu-handle x)))
  x)

;;; ---- iterators (src/iteration/linit.lisp:48-63) ----------------------------

(defmethod make-iterator ((jac <jacobi>) (mat <gpu-sparse-matrix>))
  (make-instance '<iterator> :matrix mat :residual-before t :initialize nil
    :iterate (lambda (x b r) (declare (ignore b))
               (flgpu-check (flgpu-jacobi-iterate (gpu-handle mat)
                                                  (float (slot-value jac 'damp) 1d0)
                                                  (gpu-handle x) (gpu-handle r))))
    :residual-after nil))

(defmethod make-iterator ((sor <sor>) (mat <gpu-sparse-matrix>))
  (let ((n (make-array 1 :element-type '(signed-byte 32))))
    (flgpu-check (fl.port:foreign-call #'flgpu-mat-colour-greedy (gpu-handle mat) n)))
  (make-instance '<iterator> :matrix mat :residual-before nil :initialize nil
    :iterate (lambda (x b r)
               (flgpu-check (flgpu-sor-iterate (gpu-handle mat)
                                               (float (slot-value sor 'omega) 1d0)
                                               0 (gpu-handle x) (gpu-handle b)
                                               (gpu-handle r))))
    :residual-after nil))

(defun block-iterator (iter mat kind weight)
  "<block-iteration> iterator (blockit.lisp:113-203) on the vertex-centred blocks of the
lattice numbering MAT's keys: residual-before t, residual-after nil."
  (destructuring-bind (dim order n)
      (or (gpu-lattice mat) (error "block smoothing needs the matrix's lattice"))
    (flgpu-check (flgpu-lattice-colouring (gpu-handle mat) dim order n)))
  (make-instance '<iterator> :matrix mat :residual-before t :initialize nil
    :iterate (lambda (x b r)
               (flgpu-check (flgpu-block-iterate (gpu-handle mat) kind
                                                 (float (slot-value iter weight) 1d0)
                                                 0 (gpu-handle x) (gpu-handle b)
                                                 (gpu-handle r))))
    :residual-after nil))

(defmethod make-iterator ((ssc fl.geomg::<geometric-ssc>) (mat <gpu-sparse-matrix>))
  (block-iterator ssc mat 2 'fl.iteration::omega))     ; FLGPU_SMOOTH_SSC

(defmethod make-iterator ((psc fl.geomg::<geometric-psc>) (mat <gpu-sparse-matrix>))
  (block-iterator psc mat 3 'fl.iteration::damp))      ; FLGPU_SMOOTH_PSC

;;; ---- export for offline hand-over ----------------------------------------

(defun write-le (stream value bytes)
  (dotimes (i bytes) (write-byte (ldb (byte 8 (* 8 i)) value) stream)))

(defun write-double-le (stream x)
  #+sbcl (write-le stream (logior (ash (ldb (byte 32 0) (sb-kernel:double-float-high-bits x)) 32)
                                  (sb-kernel:double-float-low-bits x))
                   8))

(defun export-system (A b path)
  "Writes the <sparse-matrix> A and the <sparse-vector> b as a system file:
columns compressed in femlisp's key order (sparse-matrix->ccs), the rhs in
that order, and every key's rank in sort-lexicographically order as the
permutation flgpu_system_read applies."
  (let* ((keys (coerce (row-keys A) 'vector))
         (rank (make-hash-table))
         (ccs (sparse-matrix->ccs A :keys keys))
         (rhs (store (sparse-vector->matlisp b keys))))
    (loop for k across (canonical-keys A) and i from 0 do (setf (gethash k rank) i))
    (with-slots (sizes starts indices) (pattern ccs)
      (with-open-file (s path :direction :output :element-type '(unsigned-byte 8)
                              :if-exists :supersede)
        (loop for c across "FLSYS01" do (write-byte (char-code c) s))
        (write-byte 0 s)
        (write-le s 1 4)                           ; columns compressed
        (write-le s 3 4)                           ; rhs and permutation follow
        (write-le s (aref sizes 1) 8)              ; rows
        (write-le s (aref sizes 0) 8)              ; columns
        (write-le s (aref starts (aref sizes 0)) 8)
        (loop for v across starts do (write-le s v 4))
        (loop for v across indices do (write-le s v 4))
        (loop for v across (store ccs) do (write-double-le s v))
        (loop for v across rhs do (write-double-le s v))
        (loop for k across keys do (write-le s (gethash k rank) 4))))
    path))
