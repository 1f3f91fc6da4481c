/*
 * pa_zmtp.h -- C ABI of libpa_zmtp.so: the ZeroMQ wire protocol (ZMTP/3.0, NULL security) for
 * PUSH/PULL endpoints, and SHA3-256.  These serve the ZeroMQ front-end (pa_server), which keeps
 * the reference's three sockets: seed-in PULL (PrivacyAmplification.cu:980-1014), key-in PULL
 * (:1036-1073) and amp-out PUSH (:1143-1200), with its frame layouts, so that the reference's
 * unchanged libzmq peers (its examples directory) can talk to it.  libzmq itself is not needed.
 */
#ifndef PA_ZMTP_H
#define PA_ZMTP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct pa_zmtp pa_zmtp;

/* role "PUSH" or "PULL"; address "tcp://host:port" ("*" binds every interface); bind != 0
 * listens for one peer, else connects (retrying until the peer is up).  NULL on bad arguments. */
pa_zmtp* pa_zmtp_open(const char* role, const char* address, int bind);
/* Sends one multipart message (nparts frames of lens[i] bytes); blocks until written.
 * Returns 0, or -1 once pa_zmtp_stop was called. */
int pa_zmtp_send(pa_zmtp* z, const void* const* parts, const uint64_t* lens, uint32_t nparts);
/* Receives one message; *nparts receives its frame count.  The frames stay readable through
 * pa_zmtp_part until the next receive.  Returns 0, or -1 once pa_zmtp_stop was called. */
int pa_zmtp_recv(pa_zmtp* z, uint32_t* nparts);
int pa_zmtp_part(pa_zmtp* z, uint32_t i, const void** data, uint64_t* len);
/* Unblocks pending and later calls (any thread); then pa_zmtp_close releases the endpoint. */
void pa_zmtp_stop(pa_zmtp* z);
void pa_zmtp_close(pa_zmtp* z);

/* SHA3-256 (FIPS 202) of data[0, len), as the reference's verify_ampout computes it
 * (PrivacyAmplification.cu:1120-1140). */
void pa_sha3_256(const void* data, uint64_t len, uint8_t digest[32]);

#ifdef __cplusplus
}
#endif
#endif
