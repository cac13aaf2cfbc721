;;;; lisp/qvm-hip.lisp
;;;;
;;;; The reference-side binding of libqvmhip (include/qvmhip.h) for qvm
;;;; 1.18.0.  qvm has no C plugin API: its seams are CLOS generics on the
;;;; pure-state class, and src/apply-gate.lisp:7-10 says how to add a state
;;;; ("If a new quantum STATE is defined, define new protocol methods for that
;;;; STATE type").  HIP-PURE-STATE is that state: a PURE-STATE subclass (so it
;;;; fits the (or null pure-state) slot of PURE-STATE-QVM, src/qvm.lisp:113-117)
;;;; whose amplitudes live in GPU HBM behind a qvmhip_state handle.  Every
;;;; protocol generic the interpreter, the compiler and the measurement code
;;;; dispatch on has a method here; everything above the protocol (LOAD-PROGRAM,
;;;; RUN, TRANSITION, the app) is the reference's own code.
;;;;
;;;; The host copy of the amplitudes is a snapshot: STATE-ELEMENTS (and so
;;;; AMPLITUDES of the state or of the QVM) copies the device state into a
;;;; cached Lisp vector, which every mutating call invalidates; (SETF
;;;; STATE-ELEMENTS) uploads.  Writes into the vector that STATE-ELEMENTS
;;;; returned reach the device only through (SETF STATE-ELEMENTS).
;;;;
;;;; Uniform draws come from qvm's own RANDOM (mt19937, src/package.lisp:10-11)
;;;; at the same points the reference draws, so a seeded run consumes the same
;;;; stream: MEASURE draws inside the reference's method (src/measurement.lisp:
;;;; 87-105) only when p1 /= 0; MEASURE-ALL draws once (:128-143); the compiled
;;;; projector always draws (src/compile-gate.lisp:221-254).
;;;;
;;;; Load after qvm (asdf: lisp/qvm-hip.asd), with libqvmhip.so on CFFI's
;;;; search path (e.g. push qvm_amd/_lib/ onto cffi:*foreign-library-directories*).

(defpackage #:qvm-hip
  (:use #:cl)
  (:local-nicknames (#:quil #:cl-quil.frontend))
  (:export
   #:hip-pure-state                     ; CLASS
   #:make-hip-pure-state                ; FUNCTION
   #:make-hip-qvm                       ; FUNCTION
   #:hip-handle                         ; READER
   #:qvmhip-error                       ; CONDITION
   #:write-hip-wavefunction             ; FUNCTION
   ))

(in-package #:qvm-hip)

;;; ------------------------------------------------------------------------
;;; The C ABI

(cffi:define-foreign-library libqvmhip
  (:unix (:or "libqvmhip.so" "qvm_amd/_lib/libqvmhip.so"))
  (t (:default "libqvmhip")))

(cffi:use-foreign-library libqvmhip)

(defconstant +f64+ 0)
(defconstant +f32+ 1)

(cffi:defcstruct qvmhip-config
  (struct-size :uint32)
  (device :int32)
  (memory-cap-bytes :uint64)
  (flags :uint32))

(cffi:defcfun ("qvmhip_last_error" %last-error) :string)
(cffi:defcfun ("qvmhip_create" %create) :int
  (n :uint32) (dtype :int) (cfg :pointer) (out :pointer))
(cffi:defcfun ("qvmhip_destroy" %destroy) :int (s :pointer))
(cffi:defcfun ("qvmhip_num_qubits" %num-qubits) :int (s :pointer) (n :pointer))
(cffi:defcfun ("qvmhip_reset_zero" %reset-zero) :int (s :pointer))
(cffi:defcfun ("qvmhip_sync" %sync) :int (s :pointer))
(cffi:defcfun ("qvmhip_apply_matrix" %apply-matrix) :int
  (s :pointer) (u :pointer) (k :uint32) (qubits :pointer))
(cffi:defcfun ("qvmhip_prob_excited" %prob-excited) :int
  (s :pointer) (q :uint32) (p1 :pointer))
(cffi:defcfun ("qvmhip_force_measure" %force-measure) :int
  (s :pointer) (q :uint32) (bit :int) (p1 :double))
(cffi:defcfun ("qvmhip_measure" %measure) :int
  (s :pointer) (q :uint32) (u :double) (compiled :int) (bit :pointer) (consumed :pointer))
(cffi:defcfun ("qvmhip_measure_all" %measure-all) :int
  (s :pointer) (u :double) (index :pointer))
(cffi:defcfun ("qvmhip_read_amplitudes" %read-amplitudes) :int
  (s :pointer) (start :uint64) (count :uint64) (out :pointer))
(cffi:defcfun ("qvmhip_write_amplitudes" %write-amplitudes) :int
  (s :pointer) (start :uint64) (count :uint64) (in :pointer))
(cffi:defcfun ("qvmhip_read_wavefunction_be" %read-wavefunction-be) :int
  (s :pointer) (start :uint64) (count :uint64) (out :pointer))

(define-condition qvmhip-error (error)
  ((code :initarg :code :reader qvmhip-error-code)
   (message :initarg :message :reader qvmhip-error-message))
  (:report (lambda (c stream)
             (format stream "qvmhip (~D): ~A" (qvmhip-error-code c) (qvmhip-error-message c))))
  (:documentation "A negative status of a qvmhip entry point, with qvmhip_last_error()'s message."))

(declaim (inline check))
(defun check (rc)
  (unless (zerop rc)
    (error 'qvmhip-error :code rc :message (%last-error)))
  rc)

(defmacro with-vector-pointer ((ptr vector) &body body)
  "Bind PTR to the address of the data of the specialised simple VECTOR
(pinned for the extent of BODY)."
  #+sbcl
  `(sb-sys:with-pinned-objects (,vector)
     (let ((,ptr (sb-sys:vector-sap ,vector)))
       ,@body))
  #-sbcl
  (declare (ignore ptr vector body))
  #-sbcl
  (error "qvm-hip pins Lisp vectors with SBCL's sb-sys; port WITH-VECTOR-POINTER to this Lisp."))

;;; ------------------------------------------------------------------------
;;; The state

(defclass hip-pure-state (qvm:pure-state)
  ((handle :initarg :handle
           :initform nil
           :reader hip-handle
           :documentation "The qvmhip_state* that owns the amplitudes in HBM.")
   (device-qubits :initarg :device-qubits
                  :reader device-qubits
                  :documentation "Number of qubits of the device state.")
   (host-buffer :initform nil
                :accessor host-buffer
                :documentation "Host vector the amplitudes are copied into (allocated on first read).")
   (snapshot-valid-p :initform nil
                     :accessor snapshot-valid-p
                     :documentation "True while HOST-BUFFER holds the device amplitudes.")
   (zero-p :initform t
           :accessor zero-p
           :documentation "True while the device state is known to be |0...0> (no copy needed to read it)."))
  (:documentation "A PURE-STATE whose 2^n amplitudes stay resident in GPU HBM. The inherited AMPLITUDES slot holds a one-element placeholder: the amplitude vector is produced on demand by STATE-ELEMENTS."))

(defun make-hip-pure-state (num-qubits &key (dtype :f64) (device -1))
  "The device counterpart of MAKE-PURE-STATE (src/state-representation.lisp:
76-102): a |0...0> state of NUM-QUBITS qubits in HBM, DTYPE :F64 (complex
double, the reference's representation) or :F32, freed by a finalizer as
MAKE-PURE-STATE's allocation finalizer frees foreign memory."
  (check-type num-qubits (integer 1 40))
  (let ((code (ecase dtype (:f64 +f64+) (:f32 +f32+))))
    (cffi:with-foreign-object (out :pointer)
      (cffi:with-foreign-object (cfg '(:struct qvmhip-config))
        (setf (cffi:foreign-slot-value cfg '(:struct qvmhip-config) 'struct-size)
              (cffi:foreign-type-size '(:struct qvmhip-config))
              (cffi:foreign-slot-value cfg '(:struct qvmhip-config) 'device) device
              (cffi:foreign-slot-value cfg '(:struct qvmhip-config) 'memory-cap-bytes) 0
              (cffi:foreign-slot-value cfg '(:struct qvmhip-config) 'flags) 0)
        (check (%create num-qubits code cfg out)))
      (let* ((h (cffi:mem-ref out :pointer))
             ;; :NUM-QUBITS 0 keeps PURE-STATE's INITIALIZE-INSTANCE :AFTER
             ;; (src/state-representation.lisp:104-124) from allocating a 2^n
             ;; host vector: it allocates a one-element placeholder instead.
             ;; The handle is set afterwards, so that method still sees the
             ;; placeholder slot (AMPLITUDES below).
             (state (make-instance 'hip-pure-state :num-qubits 0 :device-qubits num-qubits)))
        (setf (slot-value state 'handle) h)
        (tg:finalize state (lambda () (%destroy h)))
        state))))

(declaim (inline stale))
(defun stale (state)
  "Called before every call that changes the device state."
  (setf (snapshot-valid-p state) nil
        (zero-p state) nil))

;;; State protocol (src/state-representation.lisp:25-47).

(defmethod qvm::num-qubits ((state hip-pure-state))
  (device-qubits state))

(defmethod qvm:state-elements ((state hip-pure-state))
  ;; The only device-to-host copy of the path.  A state known to be
  ;; |0...0> (new, or just reset) is written on the host without a copy:
  ;; PURE-STATE-QVM's INITIALIZE-INSTANCE :AFTER reads (AMPLITUDES QVM) once
  ;; to check the qubit count (src/qvm.lisp:124-148).
  (let ((v (or (host-buffer state)
               (setf (host-buffer state)
                     (qvm::make-lisp-cflonum-vector (expt 2 (device-qubits state)))))))
    (unless (snapshot-valid-p state)
      (cond ((zero-p state)
             (fill v (qvm:cflonum 0))
             (setf (aref v 0) (qvm:cflonum 1)))
            (t
             (with-vector-pointer (p v)
               (check (%read-amplitudes (hip-handle state) 0 (length v) p)))))
      (setf (snapshot-valid-p state) t))
    v))

(defmethod (setf qvm:state-elements) (new-value (state hip-pure-state))
  (check-type new-value qvm:quantum-state)
  (assert (= (length new-value) (expt 2 (device-qubits state))) ()
          "The new amplitudes represent ~D qubit~:P, the state has ~D."
          (qvm::wavefunction-qubits new-value) (device-qubits state))
  (stale state)
  (with-vector-pointer (p new-value)
    (check (%write-amplitudes (hip-handle state) 0 (length new-value) p)))
  (setf (host-buffer state) new-value
        (snapshot-valid-p state) t)
  new-value)

;;; AMPLITUDES of a PURE-STATE is the slot accessor that the reference's
;;; pure-state code reads directly (e.g. MEASURE-ALL-STATE, the noisy
;;; machines).  Before the handle exists -- inside PURE-STATE's own
;;; INITIALIZE-INSTANCE :AFTER -- the placeholder slot is used as is.
(defmethod qvm::amplitudes ((state hip-pure-state))
  (if (hip-handle state)
      (qvm:state-elements state)
      (call-next-method)))

(defmethod (setf qvm::amplitudes) (new-value (state hip-pure-state))
  (if (hip-handle state)
      (setf (qvm:state-elements state) new-value)
      (call-next-method)))

(defmethod qvm:set-to-zero-state ((state hip-pure-state))
  (stale state)
  (check (%reset-zero (hip-handle state)))
  (setf (zero-p state) t)
  state)

(defmethod qvm::requires-swapping-amps-p ((state hip-pure-state))
  ;; No trial buffer on the device path: the amplitudes never move.
  nil)

(defmethod qvm::swap-internal-amplitude-pointers ((state hip-pure-state))
  state)

(defmethod qvm::check-allocate-computation-space ((state hip-pure-state))
  ;; The stochastic Kraus evolution (src/apply-gate.lisp:16-39) works on
  ;; host trial vectors; on this state it would evolve a stale snapshot.
  (error "HIP-PURE-STATE does not run superoperators (Kraus maps) through the host trial buffer."))

;;; Gate protocol (src/apply-gate.lisp:106-181).  Each reference method on
;;; PURE-STATE builds a QUANTUM-OPERATOR (row-major complex double,
;;; src/linear-algebra.lisp:22-47) and applies it on the NAT-TUPLE of the
;;; qubits (the last Quil argument first, src/utilities.lisp:43-51); here the
;;; same operator goes to qvmhip_apply_matrix, whose matrix convention is that
;;; of QUANTUM-OPERATOR.  A method per gate class: a method on QUIL:GATE alone
;;; would lose to the reference's (QUIL:SIMPLE-GATE PURE-STATE) methods, whose
;;; first argument is more specific.

(defun apply-operator-on-device (state operator qubit-tuple)
  "Apply the QUANTUM-OPERATOR OPERATOR (2^k x 2^k) on the nat-tuple
QUBIT-TUPLE of the device state."
  (declare (type qvm:quantum-operator operator))
  (let* ((k (length qubit-tuple))
         (storage #+sbcl (sb-ext:array-storage-vector operator)
                  #-sbcl (make-array (array-total-size operator)
                                     :element-type 'qvm:cflonum
                                     :displaced-to operator)))
    (assert (= (array-dimension operator 0) (expt 2 k)) ()
            "I attempted to apply a ~D-qubit operator to ~D qubit~:P."
            (1- (integer-length (array-dimension operator 0))) k)
    (stale state)
    (cffi:with-foreign-object (q :uint32 k)
      (dotimes (j k)
        (setf (cffi:mem-aref q :uint32 j) (aref qubit-tuple j)))
      ;; row-major (complex double-float) storage = interleaved (re, im)
      (with-vector-pointer (u storage)
        (check (%apply-matrix (hip-handle state) u k q))))
    state))

(defun gate-operator (gate parameters)
  (qvm:magicl-matrix-to-quantum-operator (apply #'quil:gate-matrix gate parameters)))

(defmacro define-device-gate-methods (&rest specs)
  `(progn
     ,@(loop :for (class no-parameters) :in specs
             :collect
             `(defmethod qvm:apply-gate-to-state ((gate ,class) (state hip-pure-state) qubits
                                                  &rest parameters)
                ,@(when no-parameters
                    '((assert (null parameters) (parameters)
                              "Parameters don't make sense for simple gates.")))
                (apply-operator-on-device state (gate-operator gate parameters)
                                          (apply #'qvm::nat-tuple qubits))))))

;;; The gate classes the reference dispatches on for pure states
;;; (src/apply-gate.lisp:109-160), with its parameter assertions.
(define-device-gate-methods
  (quil:simple-gate t)
  (quil:parameterized-gate nil)
  (quil:sequence-gate nil)
  (quil:permutation-gate t)
  (quil:controlled-gate nil)
  (quil:forked-gate nil)
  (quil:dagger-gate nil))

;;; Compiled gate applications (src/apply-gate.lisp:162-181).  Their
;;; APPLY-OPERATOR closures work on host vectors; the device path applies the
;;; compiled (static) matrix, or the permutation gate's matrix, on the
;;; application's own qubits.
(defun application-tuple (app)
  (apply #'qvm::nat-tuple (mapcar #'quil:qubit-index (quil:application-arguments app))))

(defmethod qvm:apply-gate-to-state ((gate qvm::compiled-matrix-gate-application)
                                    (state hip-pure-state) qubits &rest parameters)
  (declare (ignore qubits))
  (assert (null parameters) (parameters)
          "Parameters don't make sense for a COMPILED-MATRIX-GATE-APPLICATIONs.")
  (apply-operator-on-device state (qvm::compiled-matrix gate) (application-tuple gate)))

(defmethod qvm:apply-gate-to-state ((gate qvm::compiled-inlined-matrix-gate-application)
                                    (state hip-pure-state) qubits &rest parameters)
  (declare (ignore qubits))
  (assert (null parameters) (parameters)
          "Parameters don't make sense for a COMPILED-INLINED-MATRIX-GATE-APPLICATIONs.")
  (apply-operator-on-device state (qvm::compiled-matrix gate) (application-tuple gate)))

(defmethod qvm:apply-gate-to-state ((gate qvm::compiled-permutation-gate-application)
                                    (state hip-pure-state) qubits &rest parameters)
  (declare (ignore qubits))
  (assert (null parameters) (parameters)
          "Parameters don't make sense for a COMPILED-PERMUTATION-GATE-APPLICATIONs.")
  (apply-operator-on-device state (gate-operator (qvm::source-gate gate) nil)
                            (application-tuple gate)))

;;; Measurement protocol (src/measurement.lisp:7-143).  MEASURE itself is
;;; the reference's method on BASE-QVM: it calls the two methods below and
;;; draws in between.

(defmethod qvm::get-excited-state-probability ((state hip-pure-state) qubit)
  (cffi:with-foreign-object (p :double)
    (check (%prob-excited (hip-handle state) qubit p))
    (cffi:mem-ref p :double)))

(defmethod qvm:force-measurement (measured-value qubit (state hip-pure-state) excited-probability)
  (stale state)
  (check (%force-measure (hip-handle state) qubit measured-value
                         (coerce excited-probability 'double-float)))
  state)

(defmethod qvm::measure-all-state ((state hip-pure-state) (qvm qvm::base-qvm))
  ;; One draw, b = min{b : C(b) > p} (the bisection sampler,
  ;; src/measurement.lisp:179-225), collapse to |b> with amplitude 1.
  (stale state)
  (cffi:with-foreign-object (b :uint64)
    (check (%measure-all (hip-handle state) (qvm:flonum (qvm::random 1.0d0)) b))
    (let ((basis-state (cffi:mem-ref b :uint64)))
      (values qvm (loop :for i :below (qvm:number-of-qubits qvm)
                        :collect (ldb (byte 1 i) basis-state))))))

;;; ------------------------------------------------------------------------
;;; The machine
;;;
;;; A HIP-PURE-STATE goes into the reference's own PURE-STATE-QVM through
;;; its :STATE initarg (src/qvm.lisp:113-117, 124-148).  Two places in the
;;; machine code touch the host amplitude vector directly instead of going
;;; through the state protocol: RESET-QUANTUM-STATE (src/qvm.lisp:84-87) and
;;; the compiled projector's TRANSITION (src/transition.lisp:188-194).  The
;;; :AROUND methods below route them to the device for a HIP-PURE-STATE and
;;; leave every other state to the reference.

(defun make-hip-qvm (num-qubits &key (classical-memory-model quil:**empty-memory-model**)
                                     (dtype :f64) (device -1))
  "MAKE-QVM (src/qvm.lisp:150-164) with the state in HBM: a PURE-STATE-QVM
whose state is a HIP-PURE-STATE."
  (check-type num-qubits unsigned-byte)
  (check-type classical-memory-model quil:memory-model)
  (make-instance 'qvm:pure-state-qvm
                 :number-of-qubits num-qubits
                 :state (make-hip-pure-state num-qubits :dtype dtype :device device)
                 :classical-memory-subsystem
                 (make-instance 'qvm:classical-memory-subsystem
                                :classical-memory-model classical-memory-model)))

(defmethod qvm::reset-quantum-state :around ((qvm qvm::base-qvm))
  (let ((state (qvm::state qvm)))
    (cond ((typep state 'hip-pure-state)
           (qvm:set-to-zero-state state)
           qvm)
          (t (call-next-method)))))

(defmethod qvm:transition :around ((qvm qvm:pure-state-qvm) (instr qvm::compiled-measurement))
  ;; The compiled projector (src/compile-gate.lisp:221-254): always one
  ;; draw r; the qubit reads 0 iff r < p0; the kept half is scaled by
  ;; 1/sqrt(p_kept) and the other half zeroed.
  (let ((state (qvm::state qvm)))
    (cond
      ((typep state 'hip-pure-state)
       (let ((q (quil:qubit-index (quil:measurement-qubit instr)))
             (src (qvm::source-instruction instr)))
         (stale state)
         (cffi:with-foreign-objects ((bit :int) (consumed :int))
           (check (%measure (hip-handle state) q (qvm:flonum (qvm::random 1.0d0)) 1 bit consumed))
           (when (typep src 'quil:measure)
             (setf (qvm::dereference-mref qvm (quil:measure-address src)) (cffi:mem-ref bit :int))))
         (incf (qvm::pc qvm))
         qvm))
      (t (call-next-method)))))

;;; ------------------------------------------------------------------------
;;; The server's wavefunction reply (app/src/handle-request.lisp:132-152):
;;; big-endian octets converted on the device, chunk by chunk, in place of
;;; MAP-AMPLITUDES + WRITE-COMPLEX-DOUBLE-FLOAT-AS-BINARY.

(defun write-hip-wavefunction (state stream &key (chunk (expt 2 20)))
  "Send every amplitude of STATE as 16 big-endian octets (re, im)."
  (let ((buf (make-array (* 16 chunk) :element-type '(unsigned-byte 8)))
        (n (expt 2 (device-qubits state))))
    (loop :for start :from 0 :below n :by chunk
          :for c := (min chunk (- n start))
          :do (with-vector-pointer (p buf)
                (check (%read-wavefunction-be (hip-handle state) start c p)))
              (write-sequence buf stream :end (* 16 c)))
    (values)))
