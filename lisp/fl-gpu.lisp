;;; fl-gpu.lisp -- femlisp's side of libflgpu.so (include/flgpu.h).
;;;
;;; Load after femlisp (SBCL).  It binds the C-ABI with fl.port:def-function
;;; (src/basic/port.lisp:442-455) in the manner of femlisp's SuperLU binding
;;; (src/alien/superlu.lisp:57-75), defines <gpu-sparse-matrix> and <gpu-vector>
;;; with the matlisp generics the solver path calls (vector.lisp:91-152,224-326),
;;; make-iterator methods for <jacobi>, <sor>, <geometric-ssc>/<geometric-psc> and
;;; <geometric-cs> on GPU matrices, a SOLVE method that hands a <linear-solver> (or
;;; <safe-linear-solver>) whose matrix lives on the GPU to flgpu_solve, and
;;; EXPORT-SYSTEM, which writes a femlisp system in the layout flgpu_system_read reads.
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
  (fl.port:def-function ("flgpu_vec_norm2" flgpu-vec-norm2)
      ((x (* t)) (out (* :double))) :returning :int)
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
  (fl.port:def-function ("flgpu_mg_create_lattice" flgpu-mg-create-lattice)
      ((ctx (* t)) (dim :int) (order :int) (n0 :int) (nlevels :int) (kappa (* t)) (nk :int)
       (f :double) (asm-mode :int) (params (* t)) (out (* (* t))) (rhs (* (* t))))
    :returning :int)
  (fl.port:def-function ("flgpu_mg_destroy" flgpu-mg-destroy) ((mg (* t))) :returning :int)
  (fl.port:def-function ("flgpu_solve" flgpu-solve)
      ((mg (* t)) (a (* t)) (params (* t)) (b (* t)) (x (* t)) (report (* t))
       (defnorms (* :double)) (max-defnorms :int))
    :returning :int)
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

;;; The remaining generics the solver loop and <cg> call on vectors (vector.lisp:91-152,
;;; 224-326): their default methods walk blocks with vref, which a device vector has not.
(defmethod scalar-type ((x <gpu-vector>)) 'double-float)
(defmethod make-analog ((x <gpu-vector>)) (gpu-vector (gpu-size x)))
(defmethod copy ((x <gpu-vector>)) (copy! x (make-analog x)))
(defmethod fill! ((x <gpu-vector>) s)
  (flgpu-check (flgpu-vec-fill (gpu-handle x) (float s 1d0))) x)
(defmethod m+! ((x <gpu-vector>) (y <gpu-vector>))   ; y <- x + y
  (axpy! 1d0 x y))
(defmethod l2-norm ((x <gpu-vector>))
  (let ((out (make-array 1 :element-type 'double-float)))
    (flgpu-check (fl.port:foreign-call #'flgpu-vec-norm2 (gpu-handle x) out))
    (aref out 0)))
(defmethod norm ((x <gpu-vector>) &optional (p 2))
  (if (eql p 2)
      (l2-norm x)
      (norm (gpu-vector-to-host x) p)))

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
  ;; <sor> with a :compare by colour (linit.lisp:258-263): the lattice colouring when the
  ;; keys number a lattice (it keeps the per-cell and stencil forms), else a greedy one
  (if (gpu-lattice mat)
      (destructuring-bind (dim order n) (gpu-lattice mat)
        (flgpu-check (flgpu-lattice-colouring (gpu-handle mat) dim order n)))
      (let ((n (make-array 1 :element-type '(signed-byte 32))))
        (flgpu-check (fl.port:foreign-call #'flgpu-mat-colour-greedy (gpu-handle mat) n))))
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

;;; ---- C structs as byte vectors ---------------------------------------------
;;; The parameter and report structs of include/flgpu.h, laid out by byte offset (int 4,
;;; double 8 aligned to 8, little endian) and passed pinned like any other Lisp array.

(defparameter +mg-params-layout+
  '((smoother 0 :int) (omega 8 :double) (pre-steps 16 :int) (post-steps 20 :int)
    (base-level 24 :int) (symmetric 28 :int) (prec 32 :int) (galerkin 36 :int)
    (transfer 40 :int) (gamma 44 :int) (additive 48 :int) (fas 52 :int))
  "flgpu_mg_params: field, byte offset, type; 56 bytes")

(defparameter +solve-params-layout+
  '((outer 0 :int) (alias 4 :int) (stop 8 :int) (tol 16 :double) (maxsteps 24 :int)
    (true-residual 28 :int) (learned 32 :int) (stall-window 36 :int)
    (restart-cycle 40 :int) (fallback 44 :int))
  "flgpu_solve_params: field, byte offset, type; 48 bytes")

(defparameter +solve-report-layout+
  '((status 0 :int) (steps 4 :int) (initial-defnorm 8 :double) (final-defnorm 16 :double)
    (true-residual-norm 24 :double) (fell-back 32 :int) (seconds 40 :double)
    (fallback-used 48 :int))
  "flgpu_solve_report: field, byte offset, type; 56 bytes")

(defun pack-struct (layout size plist)
  (let ((buf (make-array size :element-type '(unsigned-byte 8) :initial-element 0)))
    (loop for (field offset type) in layout
          for value = (getf plist (intern (symbol-name field) :keyword) 0)
          do (ecase type
               (:int
                (loop for i below 4
                      do (setf (aref buf (+ offset i)) (ldb (byte 8 (* 8 i)) value))))
               (:double
                (let ((bits (logior (ash (ldb (byte 32 0)
                                              (sb-kernel:double-float-high-bits
                                               (float value 1d0)))
                                         32)
                                    (sb-kernel:double-float-low-bits (float value 1d0)))))
                  (loop for i below 8
                        do (setf (aref buf (+ offset i)) (ldb (byte 8 (* 8 i)) bits)))))))
    buf))

(defun unpack-struct (layout buf)
  (loop for (field offset type) in layout
        collect (intern (symbol-name field) :keyword)
        collect (ecase type
                  (:int (let ((u (loop for i below 4
                                       sum (ash (aref buf (+ offset i)) (* 8 i)))))
                          (if (logbitp 31 u) (- u (ash 1 32)) u)))
                  (:double (let ((u (loop for i below 8
                                          sum (ash (aref buf (+ offset i)) (* 8 i)))))
                             (sb-kernel:make-double-float
                              (let ((hi (ldb (byte 32 32) u)))
                                (if (logbitp 31 hi) (- hi (ash 1 32)) hi))
                              (ldb (byte 32 0) u)))))))

;;; ---- <geometric-cs> on the GPU (multigrid.lisp:303-393, geomg.lisp:109-142) ----------

(defun smoother-kind (smoother)
  "FLGPU_SMOOTH_* and the damping of femlisp's smoother object"
  (typecase smoother
    (fl.geomg::<geometric-ssc> (values 2 (slot-value smoother 'fl.iteration::omega)))
    (fl.geomg::<geometric-psc> (values 3 (slot-value smoother 'fl.iteration::damp)))
    (<jacobi> (values 0 (slot-value smoother 'fl.iteration::damp)))
    (<sor> (values 1 (slot-value smoother 'fl.iteration::omega)))
    (t (error "no GPU smoother for ~A" smoother))))

(defun mg-params (mgit &key (prec 0) (transfer 0) (symmetric 1))
  "flgpu_mg_params of a femlisp <geometric-cs> / <geometric-fas> object"
  (multiple-value-bind (kind omega) (smoother-kind (fl.iteration::pre-smoother mgit))
    (pack-struct +mg-params-layout+ 56
                 (list :smoother kind :omega omega
                       :pre-steps (fl.iteration::pre-steps mgit)
                       :post-steps (fl.iteration::post-steps mgit)
                       :base-level (fl.iteration::base-level mgit)
                       :symmetric symmetric :prec prec
                       :galerkin (if (fl.geomg::galerkin-p mgit) 1 0)
                       :transfer transfer :gamma (fl.iteration::gamma mgit)
                       :additive (if (eq (fl.iteration::combination-type mgit) :additive) 1 0)
                       :fas (if (typep mgit 'fl.iteration::<fas>) 1 0)))))

(defclass <gpu-hierarchy> ()
  ((handle :initarg :handle :reader gpu-handle)
   (rhs :initarg :rhs :initform nil :reader gpu-hierarchy-rhs
        :documentation "the rhs assembled with a lattice hierarchy")))

(defun gpu-hierarchy (mgit mat &key kappa-cells (n0 1) (f 1d0) (asm-mode 0) (prec 0)
                                    a-vec i-vec)
  "flgpu_mg of the <geometric-cs> MGIT for MAT.  When MAT's keys number a lattice,
(dim order n), the levels are assembled on the GPU from n0 cells per axis up to n,
with the coefficient KAPPA-CELLS (a <gpu-vector> of n^dim cell values, or NIL for
kappa = 1), as femlisp's re-discretisation does (geomg.lisp:122-126).  Otherwise
femlisp's own multilevel decomposition is uploaded: A-VEC and I-VEC are the
blackboard's :a-vec and :i-vec (multigrid.lisp:139-141), coarsest level first."
  (let ((params (mg-params mgit :prec prec)))
    (if (gpu-lattice mat)
        (destructuring-bind (dim order n) (gpu-lattice mat)
          (let ((mg-box (make-array 1 :element-type '(unsigned-byte 64)))
                (rhs-box (make-array 1 :element-type '(unsigned-byte 64)))
                (nlevels (1+ (round (log (/ n n0) 2)))))
            (flgpu-check (fl.port:foreign-call
                          #'flgpu-mg-create-lattice (flgpu-ctx) dim order n0 nlevels
                          (if kappa-cells (gpu-handle kappa-cells) (sb-sys:int-sap 0))
                          (if kappa-cells n 0) (float f 1d0) asm-mode params mg-box rhs-box))
            (make-instance '<gpu-hierarchy>
                           :handle (sb-sys:int-sap (aref mg-box 0))
                           :rhs (make-instance '<gpu-vector> :n (nrows mat)
                                               :handle (sb-sys:int-sap (aref rhs-box 0))))))
        (let* ((levels (length a-vec))
               (as (map 'vector (lambda (A) (sap-int (gpu-handle (sparse-matrix->gpu A))))
                        a-vec))
               (is (map 'vector (lambda (I)
                                  (sap-int (gpu-handle
                                            (sparse-matrix->gpu I :keys (canonical-keys I)))))
                        i-vec))
               (a-arr (make-array levels :element-type '(unsigned-byte 64)
                                         :initial-contents as))
               (i-arr (make-array (max 1 (1- levels)) :element-type '(unsigned-byte 64)
                                                       :initial-contents
                                                       (if (plusp (length is)) is '(0)))))
          (make-instance '<gpu-hierarchy>
                         :handle (flgpu-out-handle #'flgpu-mg-create (flgpu-ctx) levels
                                                   a-arr i-arr params))))))

(defvar *gpu-hierarchies* (make-hash-table :test 'eq :weakness :key)
  "the flgpu_mg built for a (mg-iteration . matrix) pair, so that <cg> and SOLVE share it")

(defun ensure-gpu-hierarchy (mgit mat)
  (let ((per-mat (or (gethash mgit *gpu-hierarchies*)
                     (setf (gethash mgit *gpu-hierarchies*)
                           (make-hash-table :test 'eq :weakness :key)))))
    (or (gethash mat per-mat)
        (setf (gethash mat per-mat) (gpu-hierarchy mgit mat)))))

(defmethod make-iterator ((mgit fl.geomg::<geometric-cs>) (mat <gpu-sparse-matrix>))
  "The <mg-iteration> iterator (multigrid.lisp:368-391) as one flgpu_mg_apply per step."
  (let ((mg (ensure-gpu-hierarchy mgit mat)))
    (make-instance '<iterator> :matrix mat :residual-before t :initialize nil
      :iterate (lambda (x b r)
                 (flgpu-check (flgpu-mg-apply (gpu-handle mg) (gpu-handle x) (gpu-handle b)
                                              (gpu-handle r))))
      :residual-after nil)))

;;; ---- solve (solve.lisp:46-58, linsolve.lisp:58-89) --------------------------------------

(defun stop-form (success-if)
  "FLGPU_STOP_* and tolerance of a success-if form (< :reduction tol) / (< :defnorm tol)"
  (when (and (consp success-if) (eq (first success-if) '<) (= (length success-if) 3))
    (case (second success-if)
      (:reduction (values 0 (third success-if)))
      (:defnorm (values 1 (third success-if))))))

(defun step-limit (failure-if)
  "maxsteps of a failure-if form (> :step n)"
  (when (and (consp failure-if) (eq (first failure-if) '>) (eq (second failure-if) :step))
    (third failure-if)))

(defun gpu-solvable-p (solver A b)
  (and (typep A '<gpu-sparse-matrix>) (typep b '<gpu-vector>)
       (stop-form (slot-value solver 'fl.iteration::success-if))
       (step-limit (slot-value solver 'fl.iteration::failure-if))
       (let ((it (slot-value solver 'iteration)))
         (or (typep it 'fl.geomg::<geometric-cs>)
             (and (typep it '<cg>)
                  (or (null (slot-value it 'fl.iteration::preconditioner))
                      (typep (slot-value it 'fl.iteration::preconditioner)
                             'fl.geomg::<geometric-cs>)))))
       (or (not (typep solver '<safe-linear-solver>))
           (typep (fallback solver) '<lu>))))

(defmethod solve :around ((solver <linear-solver>) &optional blackboard)
  "A <linear-solver> on a system whose matrix and rhs live on the GPU runs as one
flgpu_solve call (the step loop, stopping tests and, for <safe-linear-solver>, the
<lu> fallback inside the library); anything else takes femlisp's own loop, which
also works on GPU matrices and vectors through the methods above."
  (let* ((problem (getbb blackboard :problem))
         (A (if problem (matrix problem) (getbb blackboard :matrix)))
         (b (if problem (rhs problem) (getbb blackboard :rhs))))
    (if (not (gpu-solvable-p solver A b))
        (call-next-method)
        (multiple-value-bind (stop tol) (stop-form (slot-value solver 'fl.iteration::success-if))
          (let* ((it (slot-value solver 'iteration))
                 (pc (if (typep it '<cg>) (slot-value it 'fl.iteration::preconditioner) it))
                 (mg (and pc (ensure-gpu-hierarchy pc A)))
                 (maxsteps (step-limit (slot-value solver 'fl.iteration::failure-if)))
                 (params (pack-struct
                          +solve-params-layout+ 48
                          (list :outer (cond ((not (typep it '<cg>)) 0) (pc 1) (t 2))
                                :alias 1 :stop stop :tol tol :maxsteps maxsteps
                                :restart-cycle (if (typep it '<cg>)
                                                   (or (restart-cycle it) 0)
                                                   0)
                                :fallback (if (typep solver '<safe-linear-solver>) 1 0))))
                 (report (make-array 56 :element-type '(unsigned-byte 8)))
                 (defnorms (make-array (+ maxsteps 3) :element-type 'double-float))
                 (x (or (getbb blackboard :solution) (make-domain-vector-for A)))
                 (status (fl.port:foreign-call #'flgpu-solve
                                               (if mg (gpu-handle mg) (sb-sys:int-sap 0))
                                               (gpu-handle A) params (gpu-handle b)
                                               (gpu-handle x) report defnorms
                                               (length defnorms)))
                 (rep (unpack-struct +solve-report-layout+ report)))
            (when (minusp status) (flgpu-check status))
            (when (= status 2) (error "The fallback iteration has failed."))
            (setf (getbb blackboard :solution) x
                  (getbb blackboard :status) (if (zerop status) :success :failure)
                  (getbb blackboard :step) (getf rep :steps)
                  (getbb blackboard :defnorm) (getf rep :final-defnorm)
                  (getbb blackboard :initial-defnorm) (getf rep :initial-defnorm)
                  (getbb blackboard :reduction) (if (zerop (getf rep :initial-defnorm))
                                                    0d0
                                                    (/ (getf rep :final-defnorm)
                                                       (getf rep :initial-defnorm))))
            blackboard)))))

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
