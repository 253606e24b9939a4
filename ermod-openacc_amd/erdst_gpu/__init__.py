"""MI355X-native erdst hot path: host-side mirror of ERmod's engproc
(enginit / engconst / engstore) over the C ABI of lib/liberdst_hip.so."""
