"""The reference's preprocessed-data protobuf schema (crypto/proto/rlwe.proto:41-43,
crypto/proto/pir_server.proto:23-81), declared here field for field and
instantiated with the protobuf runtime, so tests can write the files
SavePirPreprocessedDataToFile / SavePirPreprocessedDataToDirectory produce
(crypto/pir/server/pir_preprocessed_data.h:52-91, 150-224) with an encoder
independent of the library under test."""
import os

import numpy as np
from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

import oracle as O

_F = descriptor_pb2.FieldDescriptorProto


def _field(msg, name, number, ftype, label=_F.LABEL_OPTIONAL, type_name=None, oneof=None):
    f = msg.field.add(name=name, number=number, type=ftype, label=label)
    if type_name:
        f.type_name = type_name
    if oneof is not None:
        f.oneof_index = oneof


def _build():
    fd = descriptor_pb2.FileDescriptorProto(name="pirgpu_test_pir_server.proto",
                                            package="pirtest", syntax="proto3")
    poly = fd.message_type.add(name="Polynomial")
    _field(poly, "encoded_coeffs", 1, _F.TYPE_BYTES)
    mat = fd.message_type.add(name="Matrix")
    mat.oneof_decl.add(name="matrix_data")
    _field(mat, "data", 1, _F.TYPE_BYTES, oneof=0)
    _field(mat, "condensed_data", 2, _F.TYPE_BYTES, oneof=0)
    _field(mat, "rows", 3, _F.TYPE_INT32)
    _field(mat, "cols", 4, _F.TYPE_INT32)
    _field(mat, "is_condensed", 5, _F.TYPE_BOOL)
    _field(mat, "original_cols", 6, _F.TYPE_INT32)
    out = fd.message_type.add(name="PreprocessMatrixPackOutput")
    _field(out, "a_tilde_agg", 1, _F.TYPE_MESSAGE, type_name=".pirtest.Polynomial")
    _field(out, "matrix", 2, _F.TYPE_MESSAGE, type_name=".pirtest.Matrix")
    _field(out, "t_vec_h", 3, _F.TYPE_MESSAGE, _F.LABEL_REPEATED, ".pirtest.Polynomial")
    data = fd.message_type.add(name="PirPreprocessedData")
    shard_outs = data.nested_type.add(name="ShardPreprocessedOutputs")
    _field(shard_outs, "outputs", 1, _F.TYPE_MESSAGE, _F.LABEL_REPEATED,
           ".pirtest.PreprocessMatrixPackOutput")
    _field(data, "db_matrices", 1, _F.TYPE_MESSAGE, _F.LABEL_REPEATED, ".pirtest.Matrix")
    _field(data, "preprocessed_outputs", 2, _F.TYPE_MESSAGE, _F.LABEL_REPEATED,
           ".pirtest.PirPreprocessedData.ShardPreprocessedOutputs")
    _field(data, "second_dim_a", 3, _F.TYPE_MESSAGE, _F.LABEL_REPEATED, ".pirtest.Polynomial")
    shard = fd.message_type.add(name="PirPreprocessedShard")
    _field(shard, "db_matrix", 1, _F.TYPE_MESSAGE, type_name=".pirtest.Matrix")
    _field(shard, "outputs", 2, _F.TYPE_MESSAGE, _F.LABEL_REPEATED,
           ".pirtest.PreprocessMatrixPackOutput")
    meta = fd.message_type.add(name="PirPreprocessedMetadata")
    _field(meta, "second_dim_a", 1, _F.TYPE_MESSAGE, _F.LABEL_REPEATED, ".pirtest.Polynomial")
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fd)
    get = lambda n: message_factory.GetMessageClass(pool.FindMessageTypeByName("pirtest." + n))  # noqa: E731
    return {n: get(n) for n in ("Polynomial", "Matrix", "PreprocessMatrixPackOutput",
                                "PirPreprocessedData", "PirPreprocessedShard",
                                "PirPreprocessedMetadata")}


M = _build()


def _poly(coeffs, logm):
    return M["Polynomial"](encoded_coeffs=O.poly_to_proto(coeffs, logm))


def _db_matrix(P, words, K):
    # Matrix::ToProto of a condensed DB matrix (crypto/matrix.cc:479-495).
    per = 4 if P.db_bytes == 2 else 8
    return M["Matrix"](condensed_data=np.ascontiguousarray(words).tobytes(), rows=words.shape[0],
                       cols=K // per, is_condensed=True, original_cols=K)


def _pack_matrix(P, rows):
    if P.mat_bytes == 4:  # condensed int32
        kd = rows.shape[1] * 2
        return M["Matrix"](condensed_data=np.ascontiguousarray(rows).tobytes(),
                           rows=rows.shape[0], cols=kd // 2, is_condensed=True,
                           original_cols=kd)
    kd = rows.shape[1]  # int16 entries: not condensed
    return M["Matrix"](data=np.ascontiguousarray(rows).tobytes(), rows=rows.shape[0], cols=kd,
                       is_condensed=False, original_cols=kd)


def _outputs(P, mat, s, shard):
    d, t, k = s.d, s.t, s.k
    logm = min(P.log_q, 64)
    outs = []
    for i in range(t):
        oi = shard * t + i
        rows = mat.pack[oi * d:(oi + 1) * d]
        outs.append(M["PreprocessMatrixPackOutput"](
            a_tilde_agg=_poly(mat.a_tilde[oi], logm), matrix=_pack_matrix(P, rows),
            t_vec_h=[_poly(mat.t_vec_h[oi, j], logm) for j in range(k)]))
    return outs


def serialize_data(P, mat) -> bytes:
    """SerializePirPreprocessedData (pir_preprocessed_data.h:52-91)."""
    s = O.sizes(P)
    logm = min(P.log_q, 64)
    msg = M["PirPreprocessedData"]()
    for shard in range(s.ell):
        msg.db_matrices.append(_db_matrix(P, mat.db[shard * s.t * s.d:(shard + 1) * s.t * s.d],
                                          mat.db_cols))
        so = msg.preprocessed_outputs.add()
        so.outputs.extend(_outputs(P, mat, s, shard))
    if s.t > 1:
        msg.second_dim_a.extend([_poly(p, logm) for p in mat.second_dim_a])
    return msg.SerializeToString()


def save_directory(P, mat, path) -> None:
    """SavePirPreprocessedDataToDirectory (pir_preprocessed_data.h:175-224)."""
    s = O.sizes(P)
    logm = min(P.log_q, 64)
    os.makedirs(path, exist_ok=True)
    meta = M["PirPreprocessedMetadata"]()
    if s.t > 1:
        meta.second_dim_a.extend([_poly(p, logm) for p in mat.second_dim_a])
    with open(os.path.join(path, "metadata.bin"), "wb") as f:
        f.write(meta.SerializeToString())
    for shard in range(s.ell):
        msg = M["PirPreprocessedShard"](
            db_matrix=_db_matrix(P, mat.db[shard * s.t * s.d:(shard + 1) * s.t * s.d],
                                 mat.db_cols),
            outputs=_outputs(P, mat, s, shard))
        with open(os.path.join(path, f"shard_{shard}.bin"), "wb") as f:
            f.write(msg.SerializeToString())
