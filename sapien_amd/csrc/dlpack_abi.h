// DLPack structures as fixed by the public DLPack ABI (data-apis DLPack,
// legacy "dltensor" capsule, versions 0.5 - 0.8). Only the layout is shared
// with other implementations: it has to match byte for byte to interoperate.
#pragma once
#include <cstdint>

namespace dlabi {

enum DeviceType : int32_t { kCPU = 1, kCUDA = 2, kCUDAHost = 3, kROCM = 10, kROCMHost = 11 };
enum TypeCode : uint8_t { kInt = 0, kUInt = 1, kFloat = 2, kBfloat = 4, kBool = 6 };

struct Device {
  int32_t device_type;
  int32_t device_id;
};
struct DataType {
  uint8_t code;
  uint8_t bits;
  uint16_t lanes;
};
struct Tensor {
  void *data;
  Device device;
  int32_t ndim;
  DataType dtype;
  int64_t *shape;
  int64_t *strides;
  uint64_t byte_offset;
};
struct ManagedTensor {
  Tensor dl_tensor;
  void *manager_ctx;
  void (*deleter)(ManagedTensor *self);
};

}  // namespace dlabi
