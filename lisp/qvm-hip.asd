;;;; lisp/qvm-hip.asd
;;;;
;;;; The device state for qvm's pure-state path: load after qvm, with
;;;; libqvmhip.so (qvm_amd/_lib/) on CFFI's foreign-library search path.

(asdf:defsystem #:qvm-hip
  :description "qvm pure states resident in MI355X HBM, through the libqvmhip C ABI."
  :license "Apache-2.0"
  :depends-on (#:qvm #:cl-quil/frontend #:cffi #:trivial-garbage)
  :serial t
  :components ((:file "qvm-hip")))
