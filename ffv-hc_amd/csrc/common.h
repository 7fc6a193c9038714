// Enumerations shared by host and device code.
#pragma once

namespace ffvhc {

// Face numbering of the reference (X_M, X_P, Y_M, Y_P, Z_M, Z_P), FFVBC / BCMTools order.
enum Face { X_M = 0, X_P = 1, Y_M = 2, Y_P = 3, Z_M = 4, Z_P = 5 };

// Kind of the block on the other side of a face.
enum FaceType { FACE_OUTER = 0, FACE_SAME = 1, FACE_COARSER = 2, FACE_FINER = 3 };

}  // namespace ffvhc
