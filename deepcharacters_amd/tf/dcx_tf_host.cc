// CPython glue behind dcx_tf_host.h (no TensorFlow dependency: built and
// tested on its own as libdcx_tf_host.so).
#include "dcx_tf_host.h"

#include <Python.h>

#include <cstring>

namespace dcx_tf {
namespace {

std::string fetch_error(const char* what) {
  std::string msg = what;
  if (PyErr_Occurred()) {
    PyObject *type = nullptr, *value = nullptr, *tb = nullptr;
    PyErr_Fetch(&type, &value, &tb);
    PyErr_NormalizeException(&type, &value, &tb);
    if (value) {
      PyObject* s = PyObject_Str(value);
      if (s) {
        const char* c = PyUnicode_AsUTF8(s);
        if (c) msg += std::string(": ") + c;
        Py_DECREF(s);
      }
    }
    Py_XDECREF(type);
    Py_XDECREF(value);
    Py_XDECREF(tb);
  }
  return msg;
}

// float32 ('f') and int32 ('i' / 'l' with itemsize 4) buffers, C-contiguous
char buffer_kind(const Py_buffer& v) {
  if (v.itemsize != 4 || !v.format) return 0;
  const char* f = v.format;
  while (*f == '<' || *f == '=' || *f == '@') ++f;
  if (f[0] == 'f' && f[1] == 0) return 'f';
  if ((f[0] == 'i' || f[0] == 'l') && f[1] == 0) return 'i';
  return 0;
}

}  // namespace

std::string CallCharacterIO(const std::string& fn, const std::vector<std::string>& args, HostTables* out) {
  if (!Py_IsInitialized()) return "the shim's file readers need the Python interpreter that loaded the op library";
  PyGILState_STATE gil = PyGILState_Ensure();
  std::string err;
  PyObject *mod = nullptr, *func = nullptr, *targs = nullptr, *res = nullptr;
  mod = PyImport_ImportModule("deepcharacters_amd.tf.character_io");
  if (!mod) err = fetch_error("import deepcharacters_amd.tf.character_io");
  if (err.empty() && !(func = PyObject_GetAttrString(mod, fn.c_str()))) err = fetch_error(fn.c_str());
  if (err.empty()) {
    targs = PyTuple_New((Py_ssize_t)args.size());
    for (size_t k = 0; k < args.size(); ++k) PyTuple_SetItem(targs, (Py_ssize_t)k, PyUnicode_FromString(args[k].c_str()));
    res = PyObject_CallObject(func, targs);
    if (!res) err = fetch_error(fn.c_str());
    else if (!PyDict_Check(res)) err = fn + " did not return a dict";
  }
  if (err.empty()) {
    PyObject *key = nullptr, *val = nullptr;
    Py_ssize_t pos = 0;
    while (err.empty() && PyDict_Next(res, &pos, &key, &val)) {
      const char* name = PyUnicode_AsUTF8(key);
      if (!name) { err = fetch_error("table name"); break; }
      Py_buffer view;
      if (PyObject_GetBuffer(val, &view, PyBUF_C_CONTIGUOUS | PyBUF_FORMAT | PyBUF_ND) != 0) {
        err = fetch_error(name);
        break;
      }
      HostArray a;
      a.kind = buffer_kind(view);
      if (!a.kind) {
        err = std::string(name) + ": not a float32 or int32 array";
      } else {
        a.bytes.resize((size_t)view.len);
        if (view.len) std::memcpy(a.bytes.data(), view.buf, (size_t)view.len);
        for (int d = 0; d < view.ndim; ++d) a.shape.push_back((long long)view.shape[d]);
        (*out)[name] = std::move(a);
      }
      PyBuffer_Release(&view);
    }
  }
  Py_XDECREF(res);
  Py_XDECREF(targs);
  Py_XDECREF(func);
  Py_XDECREF(mod);
  PyGILState_Release(gil);
  return err;
}

}  // namespace dcx_tf

// A C entry point for testing the glue without TensorFlow: runs
// CallCharacterIO(fn, {arg0[, arg1]}) and writes "name:kind:d0xd1...:checksum;"
// per table (sorted by name) into `summary`.  Returns 0, or -1 with the error
// text in `summary`.
extern "C" int dcx_tf_host_selftest(const char* fn, const char* arg0, const char* arg1, char* summary,
                                    int summary_len) {
  std::vector<std::string> args;
  if (arg0) args.push_back(arg0);
  if (arg1) args.push_back(arg1);
  dcx_tf::HostTables t;
  std::string err = dcx_tf::CallCharacterIO(fn, args, &t);
  std::string s;
  if (err.empty()) {
    for (const auto& kv : t) {
      s += kv.first + ":" + kv.second.kind + ":";
      for (size_t d = 0; d < kv.second.shape.size(); ++d) s += (d ? "x" : "") + std::to_string(kv.second.shape[d]);
      double sum = 0.0;
      for (long long k = 0; k < kv.second.size(); ++k)
        sum += kv.second.kind == 'f' ? (double)kv.second.f()[k] : (double)kv.second.i()[k] * (double)(k % 7 + 1);
      char buf[64];
      std::snprintf(buf, sizeof(buf), ":%.9e;", sum);
      s += buf;
    }
  } else {
    s = err;
  }
  std::snprintf(summary, (size_t)summary_len, "%s", s.c_str());
  return err.empty() ? 0 : -1;
}
